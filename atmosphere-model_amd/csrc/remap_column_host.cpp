// remap_column_host.cpp -- host build of device numerics for CPU-side
// checks; not used by the GPU library: the remap column (remap_column.hpp)
// and pow_kappa (atm_common.hpp).
//
// atmc_remap_column has the argument convention of remap_column
// (src/mod_remap.f90:165): var(1:nlay) on source faces coord(1:nlay+1)
// remapped onto target faces; arrays are 0-based here.
#include "remap_column.hpp"

#include <vector>

namespace {
struct ArrayCol {
    const double *x, *d, *v;
    double X(int k) const { return x[k - 1]; }
    double D(int k) const { return d[k - 1]; }
    double V(int k) const { return v[k - 1]; }
    using XR = double;
    XR xr(int k) const { return X(k); }
    XR dr(int k) const { return D(k); }
    double xv(XR r) const { return r; }
};
}  // namespace

extern "C" void atmc_remap_column(int nlay, const double *src_var, double *dst_var, const double *src_coord,
                                  const double *dst_coord) {
    std::vector<double> lv(nlay + 2), gm(nlay + 2);
    ArrayCol c{src_coord, dst_coord, src_var};
    ppm_forward(c, nlay, lv.data(), gm.data(), 1);
    auto out = [&](int m, double acc, double, double) { dst_var[m - 1] = acc; };
    ppm_merge(c, nlay, lv.data(), gm.data(), 1, out);
}

// pow_kappa on n values (atm_common.hpp), for the CPU check of its accuracy;
// width 4 is the grouped form the column scans use
extern "C" void atmc_pow_kappa(const double *x, double *y, long n, int width) {
    long i = 0;
    if (width == 4)
        for (; i + 4 <= n; i += 4) {
            double v[4] = {x[i], x[i + 1], x[i + 2], x[i + 3]};
            pow_kappa_n<4>(v);
            for (int k = 0; k < 4; k++) y[i + k] = v[k];
        }
    for (; i < n; i++) y[i] = pow_kappa(x[i]);
}
