// remap_column_host.cpp -- host build of the device remap column
// (remap_column.hpp) for CPU-side checks; not used by the GPU library.
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
