// k_output.hip -- time-mean output accumulation on the device (gfx950).
//
// src/mod_output.f90:37-160 without the netCDF writes: output_field
// (isd:ied, jsd:jed, nlay+1, 6) holds the dt-weighted sums of dp, pt, the
// A-grid u and v, ts and the net flux; write_output divides by the
// accumulated time and resets.  compensated_sum (132-142) never stores its
// compensation term (Q9), so each element is a plain running sum; the sum
// of an element is independent of every other, so one thread per element.
#include "atm_common.hpp"

namespace {

// accumulate_output 94-130: element (i, j, k) of the six fields
__global__ void k_accumulate(Geo g, double dt, const double *__restrict__ dp, const double *__restrict__ pt,
                             const double *__restrict__ ud, const double *__restrict__ vd,
                             const double *__restrict__ ts, const double *__restrict__ nf, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;  // 0 .. nlay
    if (i >= g.nx) return;
    const int L = g.L;
    const size_t fstride = (size_t)g.nx * g.ny * (L + 1);
    const size_t o = ((size_t)k * g.ny + j) * g.nx + i;
    const int I = ATM_HALO + i, J = ATM_HALO + j;
    auto add = [&](int f, double val) {
        const double y = val - 0.;  // compensation(i, j, k, idx) stays 0
        out[f * fstride + o] = out[f * fstride + o] + y;
    };
    if (k == 0) add(4, dt * ts[g.ix2(I, J)]);
    if (k < L) {
        const size_t x = g.ix(I, J, k);
        add(0, dt * dp[x]);
        add(1, dt * pt[x]);
        add(2, dt * .5 * (ud[x] + ud[x + g.SI]));
        add(3, dt * .5 * (vd[x] + vd[x + 1]));
    }
    add(5, dt * nf[g.ix(I, J, k)]);
}

// write_output 146: output_field / accumulation_time
__global__ void k_mean(double *__restrict__ out, size_t n, double time) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = out[t] / time;
}

__global__ void k_zero(double *__restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = 0.;
}

}  // namespace

static size_t out_count(const Geo &g) { return (size_t)g.nx * g.ny * (g.L + 1) * 6; }

int atm_output_alloc(atm_ctx *c) {
    const size_t n = out_count(c->g);
    if (!c->outf) ATM_HIP(c, hipMalloc(&c->outf, sizeof(double) * n));
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->outf, n);
    ATM_HIP(c, hipGetLastError());
    c->acc_time = 0.;
    return ATM_OK;
}

int atm_output_accumulate(atm_ctx *c, double dt) {
    const Geo &g = c->g;
    hipLaunchKernelGGL(k_accumulate, dim3((g.nx + 63) / 64, g.ny, g.L + 1), dim3(64), 0, c->stream, g, dt, c->dp,
                       c->pt, c->ud, c->vd, c->ts, c->net_flux, c->outf);
    ATM_HIP(c, hipGetLastError());
    // while recording a graph the time is added by each launch of it
    if (c->capturing)
        c->cap_dts.push_back(dt);
    else
        c->acc_time = c->acc_time + dt;
    return ATM_OK;
}

int atm_output_take_mean(atm_ctx *c, double *host) {
    const size_t n = out_count(c->g);
    hipLaunchKernelGGL(k_mean, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->outf, n, c->acc_time);
    ATM_HIP(c, hipMemcpyAsync(host, c->outf, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->outf, n);
    ATM_HIP(c, hipGetLastError());
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    c->acc_time = 0.;
    return ATM_OK;
}
