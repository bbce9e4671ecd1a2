// Reader for snarkjs Groth16 proving keys (.zkey, version 1), the on-disk form of the bases a
// prover's MSMs use.
//
// No counterpart in the reference, which leaves the Circom/snarkjs front end to another project
// (README.md:224-225) and only accepts bases "already in Montgomery form with R = 2^256 (the form
// snarkjs/ffjavascript keep points in)" (src/msm.rs:190-197; tests/properties.rs:221-246;
// tests/g2.rs:89-111). The layout read here is the one snarkjs writes (its zkey_utils module on
// the iden3 binary-file container):
//   "zkey", u32 version = 1, u32 section count, then sections: u32 id, u64 byte size, data.
//   1 header:       u32 protocol (1 = Groth16)
//   2 Groth16:      u32 n8q, q (n8q bytes), u32 n8r, r (n8r bytes), u32 nVars, u32 nPublic,
//                   u32 domainSize, alpha1, beta1 (G1), beta2, gamma2 (G2), delta1 (G1), delta2 (G2)
//   3 IC:           nPublic + 1 G1 points
//   4 coefficients: the constraint matrices (not needed for the MSMs)
//   5 A, 6 B1:      nVars G1 points;  7 B2: nVars G2 points
//   8 C:            nVars - nPublic - 1 G1 points;  9 H: domainSize G1 points
// Integers are little-endian. Points are affine, every coordinate 32 bytes little-endian in
// Montgomery form with R = 2^256 (the library's own device form, so sections upload as they
// are); G2 coordinates are (c0, c1) pairs in the order x.c0, x.c1, y.c0, y.c1; the point at
// infinity is all zeros (also zeros in Montgomery form).
#pragma once
#include <stdint.h>

#include <string>

namespace zkg {

enum ZkeySectionId : int {
    ZKEY_IC = 3,
    ZKEY_A = 5,
    ZKEY_B1 = 6,
    ZKEY_B2 = 7,
    ZKEY_C = 8,
    ZKEY_H = 9,
};

class Zkey {
  public:
    // Maps the file and checks its header, the Groth16 parameters (BN254 q and r, 32-byte
    // elements) and the byte size of every point section.
    explicit Zkey(const std::string& path);
    ~Zkey();
    Zkey(const Zkey&) = delete;
    Zkey& operator=(const Zkey&) = delete;

    uint32_t n_vars = 0, n_public = 0, domain_size = 0;
    // Verification-key points as stored (Montgomery, R = 2^256): G1 16 words, G2 32 words.
    uint32_t alpha1[16], beta1[16], beta2[32], gamma2[32], delta1[16], delta2[32];

    // Point sections: count, words per point (16 or 32) and the stored words.
    size_t points(int id) const;
    static uint32_t point_words(int id) { return id == ZKEY_B2 ? 32 : 16; }
    const uint32_t* words(int id) const;

  private:
    const uint8_t* base_ = nullptr;
    size_t bytes_ = 0;
    uint64_t off_[16] = {}, size_[16] = {};
    bool present_[16] = {};
};

}  // namespace zkg
