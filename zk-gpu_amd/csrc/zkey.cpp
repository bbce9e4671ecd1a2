// snarkjs .zkey reader (see zkey.hpp).
#include "zkey.hpp"

#include <fcntl.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "host_math.hpp"
#include "runtime.hpp"

namespace zkg {

namespace {

uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

uint64_t rd64(const uint8_t* p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

// true if the n8 bytes at p are the little-endian encoding of the modulus m
bool is_modulus(const uint8_t* p, uint32_t n8, const host::U256& m) {
    if (n8 != 32) return false;
    uint8_t le[32];
    for (int i = 0; i < 4; ++i) memcpy(le + 8 * i, &m.l[i], 8);
    return memcmp(p, le, 32) == 0;
}

}  // namespace

Zkey::Zkey(const std::string& path) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw Error(ZKG_ERR_INVALID, "cannot open " + path + ": " + strerror(errno));
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size < 12) {
        ::close(fd);
        throw Error(ZKG_ERR_INVALID, path + " is not a .zkey file");
    }
    bytes_ = (size_t)st.st_size;
    void* m = mmap(nullptr, bytes_, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (m == MAP_FAILED) throw Error(ZKG_ERR_INVALID, "cannot map " + path);
    base_ = static_cast<const uint8_t*>(m);
    try {
        if (memcmp(base_, "zkey", 4) != 0) throw Error(ZKG_ERR_INVALID, path + " is not a .zkey file");
        if (rd32(base_ + 4) != 1)
            throw Error(ZKG_ERR_INVALID, "unsupported .zkey version " + std::to_string(rd32(base_ + 4)));
        const uint32_t nsec = rd32(base_ + 8);
        uint64_t pos = 12;
        for (uint32_t i = 0; i < nsec; ++i) {
            if (pos + 12 > bytes_) throw Error(ZKG_ERR_INVALID, "truncated .zkey section table");
            const uint32_t id = rd32(base_ + pos);
            const uint64_t sz = rd64(base_ + pos + 4);
            pos += 12;
            if (sz > bytes_ - pos) throw Error(ZKG_ERR_INVALID, "truncated .zkey section " + std::to_string(id));
            if (id < 16) {
                if (present_[id]) throw Error(ZKG_ERR_INVALID, "duplicate .zkey section " + std::to_string(id));
                present_[id] = true;
                off_[id] = pos;
                size_[id] = sz;
            }
            pos += sz;
        }
        for (int id : {1, 2, 3, 5, 6, 7, 8, 9})
            if (!present_[id]) throw Error(ZKG_ERR_INVALID, "the .zkey has no section " + std::to_string(id));
        if (size_[1] < 4 || rd32(base_ + off_[1]) != 1)
            throw Error(ZKG_ERR_INVALID, "only Groth16 .zkey files are supported");
        // Groth16 header
        const uint8_t* h = base_ + off_[2];
        const uint64_t hs = size_[2];
        const uint64_t need = 4 + 32 + 4 + 32 + 12 + 3 * 64 + 3 * 128;
        if (hs < 8 || rd32(h) != 32 || hs < need || !is_modulus(h + 4, 32, host::fq().m) || rd32(h + 36) != 32 ||
            !is_modulus(h + 40, 32, host::fr().m))
            throw Error(ZKG_ERR_INVALID, "the .zkey is not a BN254 key (q, r or element size differ)");
        n_vars = rd32(h + 72);
        n_public = rd32(h + 76);
        domain_size = rd32(h + 80);
        if (domain_size == 0 || (domain_size & (domain_size - 1)))
            throw Error(ZKG_ERR_INVALID, "the .zkey domain size is not a power of two");
        if ((uint64_t)n_public + 1 > n_vars) throw Error(ZKG_ERR_INVALID, "the .zkey has more public inputs than variables");
        const uint8_t* v = h + 84;
        memcpy(alpha1, v, 64);
        memcpy(beta1, v + 64, 64);
        memcpy(beta2, v + 128, 128);
        memcpy(gamma2, v + 256, 128);
        memcpy(delta1, v + 384, 64);
        memcpy(delta2, v + 448, 128);
        for (int id : {3, 5, 6, 7, 8, 9}) {
            const uint64_t want = (uint64_t)points(id) * 4 * point_words(id);
            if (size_[id] != want)
                throw Error(ZKG_ERR_INVALID, "the .zkey section " + std::to_string(id) + " holds " +
                                                 std::to_string(size_[id]) + " bytes; expected " + std::to_string(want));
            if (off_[id] % 4) throw Error(ZKG_ERR_INVALID, "misaligned .zkey section " + std::to_string(id));
        }
    } catch (...) {
        munmap(const_cast<uint8_t*>(base_), bytes_);
        throw;
    }
}

Zkey::~Zkey() {
    if (base_) munmap(const_cast<uint8_t*>(base_), bytes_);
}

size_t Zkey::points(int id) const {
    switch (id) {
        case ZKEY_IC: return (size_t)n_public + 1;
        case ZKEY_A:
        case ZKEY_B1:
        case ZKEY_B2: return n_vars;
        case ZKEY_C: return (size_t)n_vars - n_public - 1;
        case ZKEY_H: return domain_size;
        default: throw Error(ZKG_ERR_INVALID, "not a point section of a .zkey: " + std::to_string(id));
    }
}

const uint32_t* Zkey::words(int id) const {
    (void)points(id);  // validates the id
    return reinterpret_cast<const uint32_t*>(base_ + off_[id]);
}

}  // namespace zkg
