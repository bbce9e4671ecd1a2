// Groth16 quotient polynomial H = (A B - C) / Z_H around the NTT bundle.
//
// No counterpart in the reference (its README lists this glue as the next step, README.md:229-230).
// Inputs are the evaluations a, b, c of A, B, C on the domain H = {w^i} of size n = 2^log_n
// (natural order, plain Fr words); the output is the n coefficients of H, where Z_H(x) = x^n - 1:
//   1. A, B, C coefficients times g^i (g = 5): inverse NTTs whose last pass multiplies element i
//      by n^-1 g^i instead of n^-1 (one per-element table, no extra pass);
//   2. forward NTTs: evaluations on the coset g H;
//   3. h_i = (a_i b_i - c_i) / (g^n - 1)  (Z_H is the constant g^n - 1 on the coset);
//   4. an inverse NTT whose last pass multiplies element i by n^-1 g^-i.
// Seven transforms, the same bundle as a Groth16 prover's. When A B - C vanishes on H (a
// satisfied constraint system), H is the exact quotient and its top coefficient is zero.
#pragma once
#include "ntt.hpp"

namespace zkg {

struct GrothH {
    Device* dev = nullptr;
    uint32_t log_n = 0;
    NttPlan ntt;
    uint32_t* shift = nullptr;    // n^-1 g^i (Montgomery), i < n
    uint32_t* unshift = nullptr;  // n^-1 g^-i
    uint32_t zk[8], zk1[8];       // (g^n - 1)^-1 * 2^512 and * 2^256 mod r

    GrothH(Device* d, uint32_t log_n);
    ~GrothH();
    // a, b, c: device pointers to n elements each; on return a holds H's coefficients, b and c
    // hold the coset evaluations of B and C (clobbered).
    void run(uint32_t* a, uint32_t* b, uint32_t* c, hipStream_t stream) const;
};

}  // namespace zkg
