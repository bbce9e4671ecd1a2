// Groth16 quotient polynomial (see h_poly.hpp).
#include "h_poly.hpp"

#include "field.hpp"
#include "host_math.hpp"

namespace zkg {

// h_i = (a_i b_i - c_i) z: Montgomery products of plain values, mont(mont(a, b), z R^2) - mont(c, z R)
__global__ __launch_bounds__(256) void h_pointwise_kernel(uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                         const uint32_t* __restrict__ c, size_t n, Fr zk, Fr zk1) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fr x = fp_load<FrParams>(a + 8 * i), y = fp_load<FrParams>(b + 8 * i), z = fp_load<FrParams>(c + 8 * i);
    fp_store(a + 8 * i, ((x * y) * zk) - (z * zk1));
}

HTables::HTables(Device* d, uint32_t ln, HForm f) : dev(d), log_n(ln), form(f) {
    const host::Modulus& F = host::fr();
    const size_t n = (size_t)1 << ln;
    // coset generator: 5 (coefficient form) or a primitive 2n-th root of unity (odd coset)
    const host::U256 g = f == HForm::coefficients ? host::U256{{5, 0, 0, 0}} : host::root_of_unity(ln + 1);
    host::U256 nn = {{0, 0, 0, 0}};
    nn.l[ln / 64] = 1ull << (ln % 64);
    const host::U256 ninv = F.inverse(nn);
    const host::U256 one = {{1, 0, 0, 0}};
    host::U256 zinv = one;
    if (f == HForm::coefficients) {
        // (g^n - 1)^-1: g^n by ln squarings
        host::U256 gm = F.to_mont(g);
        for (uint32_t k = 0; k < ln; ++k) gm = F.mont(gm, gm);
        zinv = F.inverse(F.sub(F.from_mont(gm), one));
    }
    host::to_words32(zk1, F.to_mont(zinv));
    host::to_words32(zk, F.to_mont(F.to_mont(zinv)));
    device_bind(d);
    try {
        uint32_t gw[8], nw[8];
        host::to_words32(gw, g);
        host::to_words32(nw, ninv);
        ZKG_HIP(hipMalloc(&shift, 32 * n));
        ntt_geometric_table(shift, n, gw, nw, d->stream);
        if (f == HForm::coefficients) {
            uint32_t giw[8];
            host::to_words32(giw, F.inverse(g));
            ZKG_HIP(hipMalloc(&unshift, 32 * n));
            ntt_geometric_table(unshift, n, giw, nw, d->stream);
        }
    } catch (...) {
        if (shift) (void)hipFree(shift);
        if (unshift) (void)hipFree(unshift);
        throw;
    }
}

HTables::~HTables() {
    if (shift) (void)hipFree(shift);
    if (unshift) (void)hipFree(unshift);
}

void h_compute(const NttPlan& plan, uint32_t* scratch, const HTables& t, uint32_t* a, uint32_t* b, uint32_t* c,
               hipStream_t s) {
    const size_t n = (size_t)1 << t.log_n;
    for (uint32_t* p : {a, b, c}) plan.run_with(p, true, s, t.shift, scratch);   // coefficients times g^i
    for (uint32_t* p : {a, b, c}) plan.run_with(p, false, s, nullptr, scratch);  // evaluations on g H
    Fr zkf, zk1f;
    for (int i = 0; i < 8; ++i) {
        zkf.w[i] = t.zk[i];
        zk1f.w[i] = t.zk1[i];
    }
    hipLaunchKernelGGL(h_pointwise_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, a, b, c, n, zkf, zk1f);
    ZKG_HIP(hipGetLastError());
    if (t.form == HForm::coefficients) plan.run_with(a, true, s, t.unshift, scratch);  // times n^-1 g^-i
}

GrothH::GrothH(Device* d, uint32_t ln, HForm form) : dev(d), log_n(ln), ntt(d, ln), tables(d, ln, form) {}

void GrothH::run(uint32_t* a, uint32_t* b, uint32_t* c, hipStream_t s) const { h_compute(ntt, ntt.scratch, tables, a, b, c, s); }

}  // namespace zkg
