// Importance-probability files (the reference's "UAP v2" format,
// src/io/importance_probability_io.cpp) and the SHA-256 digest that binds a
// file to the representative matrix it was computed from (io/sha256.hpp).
//
// Header-only C++20. Layout (all integers little-endian):
//   bytes  0.. 7  magic "UHANNPR2"
//   byte   8      kind: 1 = L1, 2 = L2
//   bytes  9..15  reserved, zero
//   bytes 16..23  representative count (u64)
//   bytes 24..31  coordinate count d (u64)
//   bytes 32..63  SHA-256 of the representative-matrix file
//   bytes 64..    d binary64 probabilities (IEEE-754 bit patterns, u64)
// and nothing after them. Probabilities must be finite and in [0, 1].
// Errors: std::runtime_error("invalid importance-probability file <path>:
// <reason>") for I/O and format problems, std::invalid_argument for an
// unknown kind, as in the reference.
#pragma once

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace ultrahigh_ann_amd {

using Sha256Digest = std::array<std::uint8_t, 32>;

// SHA-256 (FIPS PUB 180-4, section 6.2): 512-bit blocks, 64 rounds, the
// message padded with 0x80, zeros and its bit length (big-endian u64).
class Sha256 {
public:
    void update(const std::uint8_t* data, std::size_t n)
    {
        total_ += n;
        while (n != 0) {
            const std::size_t take = std::min(n, std::size_t{64} - used_);
            std::memcpy(block_.data() + used_, data, take);
            used_ += take;
            data += take;
            n -= take;
            if (used_ == 64) {
                compress(block_.data());
                used_ = 0;
            }
        }
    }

    Sha256Digest finish()
    {
        const std::uint64_t bits = total_ * 8U;
        const std::uint8_t one = 0x80;
        update(&one, 1);
        const std::uint8_t zero = 0;
        while (used_ != 56) update(&zero, 1);
        std::array<std::uint8_t, 8> len{};
        for (int i = 0; i < 8; ++i) len[i] = static_cast<std::uint8_t>(bits >> (56 - 8 * i));
        update(len.data(), len.size());
        Sha256Digest out{};
        for (int i = 0; i < 8; ++i) {
            for (int b = 0; b < 4; ++b) out[4 * i + b] = static_cast<std::uint8_t>(h_[i] >> (24 - 8 * b));
        }
        return out;
    }

private:
    static constexpr std::array<std::uint32_t, 64> kRound{
        0x428a2f98U, 0x71374491U, 0xb5c0fbcfU, 0xe9b5dba5U, 0x3956c25bU, 0x59f111f1U, 0x923f82a4U, 0xab1c5ed5U,
        0xd807aa98U, 0x12835b01U, 0x243185beU, 0x550c7dc3U, 0x72be5d74U, 0x80deb1feU, 0x9bdc06a7U, 0xc19bf174U,
        0xe49b69c1U, 0xefbe4786U, 0x0fc19dc6U, 0x240ca1ccU, 0x2de92c6fU, 0x4a7484aaU, 0x5cb0a9dcU, 0x76f988daU,
        0x983e5152U, 0xa831c66dU, 0xb00327c8U, 0xbf597fc7U, 0xc6e00bf3U, 0xd5a79147U, 0x06ca6351U, 0x14292967U,
        0x27b70a85U, 0x2e1b2138U, 0x4d2c6dfcU, 0x53380d13U, 0x650a7354U, 0x766a0abbU, 0x81c2c92eU, 0x92722c85U,
        0xa2bfe8a1U, 0xa81a664bU, 0xc24b8b70U, 0xc76c51a3U, 0xd192e819U, 0xd6990624U, 0xf40e3585U, 0x106aa070U,
        0x19a4c116U, 0x1e376c08U, 0x2748774cU, 0x34b0bcb5U, 0x391c0cb3U, 0x4ed8aa4aU, 0x5b9cca4fU, 0x682e6ff3U,
        0x748f82eeU, 0x78a5636fU, 0x84c87814U, 0x8cc70208U, 0x90befffaU, 0xa4506cebU, 0xbef9a3f7U, 0xc67178f2U,
    };

    static std::uint32_t rotr(std::uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

    void compress(const std::uint8_t* p)
    {
        std::array<std::uint32_t, 64> w{};
        for (int t = 0; t < 16; ++t) {
            w[t] = (std::uint32_t{p[4 * t]} << 24) | (std::uint32_t{p[4 * t + 1]} << 16) |
                   (std::uint32_t{p[4 * t + 2]} << 8) | std::uint32_t{p[4 * t + 3]};
        }
        for (int t = 16; t < 64; ++t) {
            const std::uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const std::uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        std::uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
        for (int t = 0; t < 64; ++t) {
            const std::uint32_t t1 =
                h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kRound[t] + w[t];
            const std::uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            h = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h_[0] += a;
        h_[1] += b;
        h_[2] += c;
        h_[3] += d;
        h_[4] += e;
        h_[5] += f;
        h_[6] += g;
        h_[7] += h;
    }

    std::array<std::uint32_t, 8> h_{0x6a09e667U, 0xbb67ae85U, 0x3c6ef372U, 0xa54ff53aU,
                                    0x510e527fU, 0x9b05688cU, 0x1f83d9abU, 0x5be0cd19U};
    std::array<std::uint8_t, 64> block_{};
    std::size_t used_ = 0;
    std::uint64_t total_ = 0;
};

// sha256_file (io/sha256.hpp): the digest of the file's exact bytes.
inline Sha256Digest sha256_file(const std::filesystem::path& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path.string() + " for hashing");
    Sha256 hash;
    std::vector<char> buf(1 << 20);
    while (in) {
        in.read(buf.data(), static_cast<std::streamsize>(buf.size()));
        const std::streamsize got = in.gcount();
        if (got > 0) hash.update(reinterpret_cast<const std::uint8_t*>(buf.data()), static_cast<std::size_t>(got));
    }
    if (in.bad()) throw std::runtime_error("failed while hashing " + path.string());
    return hash.finish();
}

inline std::string sha256_hex(const Sha256Digest& digest)
{
    static constexpr char kHex[] = "0123456789abcdef";
    std::string s;
    for (const std::uint8_t b : digest) {
        s.push_back(kHex[b >> 4]);
        s.push_back(kHex[b & 15]);
    }
    return s;
}

enum class ProbabilityKind : std::uint8_t { l1 = 1, l2 = 2 };

struct ProbabilityFile {
    ProbabilityKind kind{};
    std::uint64_t representative_count = 0;
    Sha256Digest representatives_sha256{};
    std::vector<double> probabilities;
};

namespace probability_file_detail {

constexpr std::array<char, 8> kMagic{'U', 'H', 'A', 'N', 'N', 'P', 'R', '2'};
constexpr std::size_t kHeaderBytes = 64;

inline std::runtime_error file_error(const std::filesystem::path& path, std::string_view why)
{
    return std::runtime_error("invalid importance-probability file " + path.string() + ": " + std::string(why));
}

inline void check_values(std::span<const double> p, const std::filesystem::path& path)
{
    for (const double v : p) {
        if (!std::isfinite(v) || v < 0.0 || v > 1.0) {
            throw file_error(path, "probabilities must be finite and in [0, 1]");
        }
    }
}

inline void put_u64(std::uint8_t* out, std::uint64_t v)
{
    for (int i = 0; i < 8; ++i) out[i] = static_cast<std::uint8_t>(v >> (8 * i));
}

inline std::uint64_t get_u64(const std::uint8_t* in)
{
    std::uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= std::uint64_t{in[i]} << (8 * i);
    return v;
}

}  // namespace probability_file_detail

// save_importance_probabilities (importance_probability_io.cpp:124-162):
// values are checked before the file is touched; parent directories are
// created; an existing file is replaced.
inline void save_probabilities(const std::filesystem::path& path, ProbabilityKind kind,
                               std::uint64_t representative_count, const Sha256Digest& digest,
                               std::span<const double> probabilities)
{
    namespace pf = probability_file_detail;
    if (kind != ProbabilityKind::l1 && kind != ProbabilityKind::l2) {
        throw std::invalid_argument("unknown importance-probability kind");
    }
    pf::check_values(probabilities, path);
    std::vector<std::uint8_t> bytes(pf::kHeaderBytes + 8 * probabilities.size(), 0);
    std::memcpy(bytes.data(), pf::kMagic.data(), pf::kMagic.size());
    bytes[8] = static_cast<std::uint8_t>(kind);
    pf::put_u64(bytes.data() + 16, representative_count);
    pf::put_u64(bytes.data() + 24, probabilities.size());
    std::memcpy(bytes.data() + 32, digest.data(), digest.size());
    for (std::size_t i = 0; i < probabilities.size(); ++i) {
        pf::put_u64(bytes.data() + pf::kHeaderBytes + 8 * i, std::bit_cast<std::uint64_t>(probabilities[i]));
    }
    if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw pf::file_error(path, "cannot open for writing");
    out.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    if (!out) throw pf::file_error(path, "failed while writing");
}

// load_importance_probabilities (importance_probability_io.cpp:164-227).
inline ProbabilityFile load_probabilities(const std::filesystem::path& path)
{
    namespace pf = probability_file_detail;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw pf::file_error(path, "does not exist or cannot be opened");
    std::array<std::uint8_t, pf::kHeaderBytes> head{};
    auto read = [&](std::uint8_t* dst, std::size_t n, std::string_view what) {
        in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(n));
        if (in.gcount() != static_cast<std::streamsize>(n)) {
            throw pf::file_error(path, "truncated while reading " + std::string(what));
        }
    };
    read(head.data(), 8, "magic");
    if (std::memcmp(head.data(), pf::kMagic.data(), pf::kMagic.size()) != 0) {
        throw pf::file_error(path, "wrong magic or unsupported version");
    }
    read(head.data() + 8, 1, "probability kind");
    if (head[8] != 1 && head[8] != 2) throw pf::file_error(path, "unknown probability kind");
    read(head.data() + 9, 7, "reserved header bytes");
    for (int i = 9; i < 16; ++i) {
        if (head[i] != 0) throw pf::file_error(path, "reserved header bytes are not zero");
    }
    read(head.data() + 16, 8, "representative count");
    read(head.data() + 24, 8, "coordinate count");
    const std::uint64_t count = pf::get_u64(head.data() + 24);
    if (count > std::vector<double>{}.max_size()) throw pf::file_error(path, "stored array is too large");
    read(head.data() + 32, 32, "representative-matrix SHA-256");
    ProbabilityFile f;
    f.kind = static_cast<ProbabilityKind>(head[8]);
    f.representative_count = pf::get_u64(head.data() + 16);
    std::memcpy(f.representatives_sha256.data(), head.data() + 32, 32);
    // Read in bounded chunks so a corrupt count fails as truncation rather
    // than as one huge allocation.
    constexpr std::uint64_t kChunk = 1 << 16;
    std::vector<std::uint8_t> raw;
    for (std::uint64_t done = 0; done < count;) {
        const std::uint64_t n = std::min(kChunk, count - done);
        raw.resize(8 * n);
        read(raw.data(), raw.size(), "probability");
        for (std::uint64_t i = 0; i < n; ++i) f.probabilities.push_back(std::bit_cast<double>(pf::get_u64(raw.data() + 8 * i)));
        done += n;
    }
    if (in.peek() != std::char_traits<char>::eof()) throw pf::file_error(path, "unexpected trailing data");
    pf::check_values(f.probabilities, path);
    return f;
}

// importance_probability_source_matches (importance_probability_io.cpp:229-241).
inline bool probability_source_matches(const ProbabilityFile& f, ProbabilityKind kind,
                                       std::uint64_t representative_count, std::uint64_t coordinate_count,
                                       const Sha256Digest& digest) noexcept
{
    return f.kind == kind && f.representative_count == representative_count &&
           f.probabilities.size() == coordinate_count && f.representatives_sha256 == digest;
}

}  // namespace ultrahigh_ann_amd
