// Groth16 quotient polynomial H = (A B - C) / Z_H around the NTT bundle.
//
// No counterpart in the reference (its README lists this glue as the next step, README.md:229-230).
// Inputs are the evaluations a, b, c of A, B, C on the domain H = {w^i} of size n = 2^log_n
// (natural order, plain Fr words). Two output forms, for the two ways proving keys hold H's bases:
//  - coefficients (bases tau^i Z(tau) / delta, as arkworks-style keys): H's n coefficients, where
//    Z_H(x) = x^n - 1:
//      1. A, B, C coefficients times g^i (g = 5): inverse NTTs whose last pass multiplies element
//         i by n^-1 g^i instead of n^-1 (one per-element table, no extra pass);
//      2. forward NTTs: evaluations on the coset g H;
//      3. h_i = (a_i b_i - c_i) / (g^n - 1)  (Z_H is the constant g^n - 1 on the coset);
//      4. an inverse NTT whose last pass multiplies element i by n^-1 g^-i.
//    When A B - C vanishes on H (a satisfied system), H is the exact quotient of degree <= n - 2.
//  - odd coset (snarkjs .zkey keys, whose H bases absorb the division): steps 1-2 with
//    g = w_2n (a primitive 2n-th root of unity, so g H is the odd 2n-th roots), then
//    a_i b_i - c_i; these n evaluations are the H MSM's scalars (snarkjs's prover).
#pragma once
#include <memory>

#include "ntt.hpp"

namespace zkg {

enum class HForm : int { coefficients = 0, odd_coset = 1 };

// The per-element scale tables and the constant of one size and form (read-only, shareable).
struct HTables {
    Device* dev = nullptr;
    uint32_t log_n = 0;
    HForm form = HForm::coefficients;
    uint32_t* shift = nullptr;    // n^-1 g^i (Montgomery), i < n
    uint32_t* unshift = nullptr;  // n^-1 g^-i (coefficient form only)
    uint32_t zk[8], zk1[8];       // z * 2^512 and z * 2^256 mod r, z = (g^n - 1)^-1 or 1
    HTables(Device* d, uint32_t log_n, HForm form);
    ~HTables();
    HTables(const HTables&) = delete;
    HTables& operator=(const HTables&) = delete;
};

// The seven (coefficients) or six (odd coset) transforms and the pointwise step on `stream`:
// a receives the output, b and c are clobbered. `plan` gives the twiddles of 2^log_n, `scratch`
// n elements.
void h_compute(const NttPlan& plan, uint32_t* scratch, const HTables& t, uint32_t* a, uint32_t* b, uint32_t* c,
               hipStream_t stream);

struct GrothH {
    Device* dev = nullptr;
    uint32_t log_n = 0;
    NttPlan ntt;
    HTables tables;

    GrothH(Device* d, uint32_t log_n, HForm form = HForm::coefficients);
    // a, b, c: device pointers to n elements each; on return a holds H's coefficients (or the odd
    // coset values), b and c hold the coset evaluations of B and C (clobbered).
    void run(uint32_t* a, uint32_t* b, uint32_t* c, hipStream_t stream) const;
};

}  // namespace zkg
