// FLAC (RFC 9639) writer and reader for PCM16 output, replacing the
// libsndfile/libFLAC calls of src/io/audio.cpp:236-262 and :290-316.
// The encoder uses fixed-blocksize frames of 4096 samples, the best of the
// constant, verbatim and fixed-predictor (order 0-4) subframes, Rice-coded
// residuals with a partition-order search, and the best stereo decorrelation
// (independent, left/side, right/side, mid/side) from level 1 up.
#include "audio.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <stdexcept>

namespace clay {
namespace {

constexpr int kBlock = 4096;

class BitWriter {
public:
    void put(uint64_t value, int bits)
    {
        for (int i = bits - 1; i >= 0; --i) {
            acc_ = (acc_ << 1) | ((value >> i) & 1);
            if (++fill_ == 8) {
                bytes.push_back(uint8_t(acc_));
                acc_ = 0;
                fill_ = 0;
            }
        }
    }
    void putSigned(int64_t value, int bits) { put(uint64_t(value) & ((bits == 64 ? 0 : (1ull << bits)) - 1), bits); }
    void unary(uint32_t zeros)
    {
        while (zeros >= 32) {
            put(0, 32);
            zeros -= 32;
        }
        put(1, int(zeros) + 1);
    }
    void align()
    {
        if (fill_)
            put(0, 8 - fill_);
    }
    std::vector<uint8_t> bytes;

private:
    uint32_t acc_ = 0;
    int fill_ = 0;
};

uint8_t crc8(const uint8_t* p, size_t n)
{
    uint8_t crc = 0;
    for (size_t i = 0; i < n; ++i) {
        crc ^= p[i];
        for (int b = 0; b < 8; ++b)
            crc = uint8_t((crc & 0x80) ? (crc << 1) ^ 0x07 : crc << 1);
    }
    return crc;
}

uint16_t crc16(const uint8_t* p, size_t n)
{
    uint16_t crc = 0;
    for (size_t i = 0; i < n; ++i) {
        crc ^= uint16_t(p[i] << 8);
        for (int b = 0; b < 8; ++b)
            crc = uint16_t((crc & 0x8000) ? (crc << 1) ^ 0x8005 : crc << 1);
    }
    return crc;
}

// Residual of the fixed predictor of `order` (RFC 9639 section 9.2.5).
void fixedResidual(const std::vector<int64_t>& x, int order, std::vector<int64_t>& r)
{
    r.resize(x.size());
    for (size_t i = size_t(order); i < x.size(); ++i) {
        switch (order) {
        case 0: r[i] = x[i]; break;
        case 1: r[i] = x[i] - x[i - 1]; break;
        case 2: r[i] = x[i] - 2 * x[i - 1] + x[i - 2]; break;
        case 3: r[i] = x[i] - 3 * x[i - 1] + 3 * x[i - 2] - x[i - 3]; break;
        default: r[i] = x[i] - 4 * x[i - 1] + 6 * x[i - 2] - 4 * x[i - 3] + x[i - 4]; break;
        }
    }
}

uint64_t zigzag(int64_t v) { return v >= 0 ? uint64_t(v) << 1 : (uint64_t(-v) << 1) - 1; }

// Best Rice parameter and its bit count for residual[first, end).
std::pair<int, uint64_t> riceCost(const std::vector<int64_t>& r, size_t first, size_t end)
{
    uint64_t sum = 0;
    for (size_t i = first; i < end; ++i)
        sum += zigzag(r[i]);
    const size_t n = end - first;
    int best = 0;
    uint64_t bestBits = std::numeric_limits<uint64_t>::max();
    for (int k = 0; k <= 14; ++k) {
        // Exact cost: sum of (u >> k) + n * (k + 1).
        uint64_t bits = uint64_t(n) * uint64_t(k + 1);
        for (size_t i = first; i < end; ++i)
            bits += zigzag(r[i]) >> k;
        if (bits < bestBits) {
            bestBits = bits;
            best = k;
        }
        if (n && (sum >> k) < n)
            break;
    }
    return {best, bestBits};
}

struct Residual {
    int order = 0;       // partition order
    std::vector<int> params;
    uint64_t bits = std::numeric_limits<uint64_t>::max();
};

Residual bestResidual(const std::vector<int64_t>& r, size_t n, int predictor, int maxOrder)
{
    Residual best;
    for (int po = 0; po <= maxOrder; ++po) {
        if ((n >> po) << po != n || (n >> po) <= size_t(predictor))
            break;
        Residual cand;
        cand.order = po;
        cand.bits = 6;  // method + partition order
        const size_t part = n >> po;
        for (int p = 0; p < (1 << po); ++p) {
            const size_t first = p == 0 ? size_t(predictor) : size_t(p) * part;
            const auto [k, bits] = riceCost(r, first, size_t(p + 1) * part);
            cand.params.push_back(k);
            cand.bits += 4 + bits;
        }
        if (cand.bits < best.bits)
            best = cand;
    }
    return best;
}

struct Subframe {
    int kind = 1;        // 0 constant, 1 verbatim, 2 fixed
    int order = 0;
    Residual residual;
    uint64_t bits = 0;
};

Subframe chooseSubframe(const std::vector<int64_t>& x, int bps, int maxPartition, std::vector<int64_t>& best)
{
    Subframe out;
    const size_t n = x.size();
    out.kind = 1;
    out.bits = 8 + uint64_t(n) * uint64_t(bps);
    if (std::all_of(x.begin(), x.end(), [&](int64_t v) { return v == x[0]; })) {
        out.kind = 0;
        out.bits = 8 + uint64_t(bps);
        return out;
    }
    std::vector<int64_t> r;
    for (int order = 0; order <= 4 && size_t(order) < n; ++order) {
        fixedResidual(x, order, r);
        const auto res = bestResidual(r, n, order, maxPartition);
        if (res.params.empty())
            continue;
        const uint64_t bits = 8 + uint64_t(order) * uint64_t(bps) + res.bits;
        if (bits < out.bits) {
            out.kind = 2;
            out.order = order;
            out.residual = res;
            out.bits = bits;
            best = r;
        }
    }
    return out;
}

void writeSubframe(BitWriter& w, const std::vector<int64_t>& x, int bps, const Subframe& s, const std::vector<int64_t>& r)
{
    if (s.kind == 0) {
        w.put(0, 8);
        w.putSigned(x[0], bps);
        return;
    }
    if (s.kind == 1) {
        w.put(0x02, 8);  // 0 + 000001 + 0
        for (auto v : x)
            w.putSigned(v, bps);
        return;
    }
    w.put(uint64_t(0x08 | s.order) << 1, 8);  // 0 + 001ooo + 0
    for (int i = 0; i < s.order; ++i)
        w.putSigned(x[size_t(i)], bps);
    w.put(0, 2);
    w.put(uint64_t(s.residual.order), 4);
    const size_t part = x.size() >> s.residual.order;
    for (int p = 0; p < (1 << s.residual.order); ++p) {
        const int k = s.residual.params[size_t(p)];
        w.put(uint64_t(k), 4);
        const size_t first = p == 0 ? size_t(s.order) : size_t(p) * part;
        for (size_t i = first; i < size_t(p + 1) * part; ++i) {
            const uint64_t u = zigzag(r[i]);
            w.unary(uint32_t(u >> k));
            if (k)
                w.put(u & ((1ull << k) - 1), k);
        }
    }
}

void putUtf8(BitWriter& w, uint64_t v)
{
    if (v < 0x80) {
        w.put(v, 8);
        return;
    }
    int extra = 1;
    while (v >= (1ull << (5 * extra + 6)))
        ++extra;
    const int leadBits = 6 - extra;
    w.put(((0xff00u >> (extra + 1)) & 0xff) | (v >> (6 * extra)), 8);
    (void)leadBits;
    for (int i = extra - 1; i >= 0; --i)
        w.put(0x80 | ((v >> (6 * i)) & 0x3f), 8);
}

class BitReader {
public:
    BitReader(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    uint64_t get(int bits)
    {
        uint64_t v = 0;
        for (int i = 0; i < bits; ++i) {
            if (pos_ >= n_ * 8)
                throw std::runtime_error("truncated FLAC");
            v = (v << 1) | ((p_[pos_ >> 3] >> (7 - (pos_ & 7))) & 1);
            ++pos_;
        }
        return v;
    }
    int64_t getSigned(int bits)
    {
        const uint64_t v = get(bits);
        return bits && (v >> (bits - 1)) ? int64_t(v) - (int64_t(1) << bits) : int64_t(v);
    }
    uint32_t unary()
    {
        uint32_t zeros = 0;
        while (!get(1))
            ++zeros;
        return zeros;
    }
    void align() { pos_ = (pos_ + 7) & ~size_t(7); }
    size_t byte() const { return pos_ >> 3; }

private:
    const uint8_t* p_;
    size_t n_, pos_ = 0;
};

void readResidual(BitReader& br, int predictor, size_t n, std::vector<int64_t>& out)
{
    const int method = int(br.get(2));
    if (method > 1)
        throw std::runtime_error("bad FLAC residual method");
    const int paramBits = method ? 5 : 4, escape = method ? 31 : 15;
    const int order = int(br.get(4));
    const size_t part = n >> order;
    for (int p = 0; p < (1 << order); ++p) {
        const int k = int(br.get(paramBits));
        const size_t count = p == 0 ? part - size_t(predictor) : part;
        if (k == escape) {
            const int bits = int(br.get(5));
            for (size_t i = 0; i < count; ++i)
                out.push_back(bits ? br.getSigned(bits) : 0);
            continue;
        }
        for (size_t i = 0; i < count; ++i) {
            const uint64_t u = (uint64_t(br.unary()) << k) | br.get(k);
            out.push_back(u & 1 ? -int64_t(u >> 1) - 1 : int64_t(u >> 1));
        }
    }
}

void readSubframe(BitReader& br, size_t n, int bps, std::vector<int64_t>& x)
{
    if (br.get(1))
        throw std::runtime_error("bad FLAC subframe");
    const int type = int(br.get(6));
    int wasted = 0;
    if (br.get(1)) {
        wasted = 1;
        while (!br.get(1))
            ++wasted;
    }
    bps -= wasted;
    x.clear();
    if (type == 0) {
        x.assign(n, br.getSigned(bps));
    } else if (type == 1) {
        for (size_t i = 0; i < n; ++i)
            x.push_back(br.getSigned(bps));
    } else if (type >= 8 && type <= 12) {
        const int order = type - 8;
        for (int i = 0; i < order; ++i)
            x.push_back(br.getSigned(bps));
        std::vector<int64_t> r;
        readResidual(br, order, n, r);
        for (size_t i = 0; i < r.size(); ++i) {
            const size_t k = x.size();
            int64_t pred = 0;
            switch (order) {
            case 1: pred = x[k - 1]; break;
            case 2: pred = 2 * x[k - 1] - x[k - 2]; break;
            case 3: pred = 3 * x[k - 1] - 3 * x[k - 2] + x[k - 3]; break;
            case 4: pred = 4 * x[k - 1] - 6 * x[k - 2] + 4 * x[k - 3] - x[k - 4]; break;
            default: break;
            }
            x.push_back(pred + r[i]);
        }
    } else if (type >= 32) {
        const int order = type - 31;
        for (int i = 0; i < order; ++i)
            x.push_back(br.getSigned(bps));
        const int precision = int(br.get(4)) + 1;
        const int shift = int(br.getSigned(5));
        std::vector<int64_t> coef(static_cast<size_t>(order), 0);
        for (auto& c : coef)
            c = br.getSigned(precision);
        std::vector<int64_t> r;
        readResidual(br, order, n, r);
        for (size_t i = 0; i < r.size(); ++i) {
            const size_t k = x.size();
            int64_t sum = 0;
            for (int j = 0; j < order; ++j)
                sum += coef[size_t(j)] * x[k - 1 - size_t(j)];
            x.push_back((shift >= 0 ? sum >> shift : sum << -shift) + r[i]);
        }
    } else {
        throw std::runtime_error("reserved FLAC subframe type");
    }
    if (wasted)
        for (auto& v : x)
            v <<= wasted;
}

} // namespace

// MD5 message digest (R. Rivest, RFC 1321) for the STREAMINFO signature of
// the unencoded audio: 64-byte blocks, four rounds of sixteen steps with the
// sine-derived constants T[i] = floor(2^32 |sin(i + 1)|).
class Md5 {
public:
    Md5()
    {
        for (int i = 0; i < 64; ++i)
            t_[i] = uint32_t(std::floor(std::fabs(std::sin(double(i + 1))) * 4294967296.0));
    }
    void update(const uint8_t* p, size_t n)
    {
        length_ += n;
        while (n) {
            const size_t take = std::min(n, size_t(64) - used_);
            std::memcpy(buffer_ + used_, p, take);
            used_ += take;
            p += take;
            n -= take;
            if (used_ == 64) {
                block(buffer_);
                used_ = 0;
            }
        }
    }
    void finish(uint8_t out[16])
    {
        const uint64_t bits = length_ * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (used_ != 56)
            update(&zero, 1);
        uint8_t tail[8];
        for (int i = 0; i < 8; ++i)
            tail[i] = uint8_t(bits >> (8 * i));
        update(tail, 8);
        const uint32_t words[4] = {a_, b_, c_, d_};
        for (int i = 0; i < 16; ++i)
            out[i] = uint8_t(words[i / 4] >> (8 * (i % 4)));
    }

private:
    static uint32_t rotate(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }
    void block(const uint8_t* p)
    {
        static const int shifts[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
        uint32_t m[16];
        for (int i = 0; i < 16; ++i)
            m[i] = uint32_t(p[4 * i]) | uint32_t(p[4 * i + 1]) << 8 | uint32_t(p[4 * i + 2]) << 16 |
                   uint32_t(p[4 * i + 3]) << 24;
        uint32_t a = a_, b = b_, c = c_, d = d_;
        for (int i = 0; i < 64; ++i) {
            const int round = i / 16;
            uint32_t f;
            int g;
            if (round == 0) { f = (b & c) | (~b & d); g = i; }
            else if (round == 1) { f = (b & d) | (c & ~d); g = (5 * i + 1) % 16; }
            else if (round == 2) { f = b ^ c ^ d; g = (3 * i + 5) % 16; }
            else { f = c ^ (b | ~d); g = (7 * i) % 16; }
            const uint32_t next = b + rotate(a + f + t_[i] + m[g], shifts[round][i % 4]);
            a = d;
            d = c;
            c = b;
            b = next;
        }
        a_ += a;
        b_ += b;
        c_ += c;
        d_ += d;
    }
    uint32_t t_[64];
    uint32_t a_ = 0x67452301u, b_ = 0xefcdab89u, c_ = 0x98badcfeu, d_ = 0x10325476u;
    uint64_t length_ = 0;
    uint8_t buffer_[64];
    size_t used_ = 0;
};

void writeFlac(const std::string& path, const int16_t* pcm, int dataFrames, int frames, int rate, int channels, int level)
{
    if (frames <= 0 || dataFrames < 0 || dataFrames > frames || (channels != 1 && channels != 2) || level < 0 || level > 8)
        throw std::runtime_error("cannot write FLAC " + path);
    const int maxPartition = std::min(8, 3 + level / 2);
    BitWriter out;
    out.bytes.insert(out.bytes.end(), {'f', 'L', 'a', 'C'});
    out.put(1, 1);
    out.put(0, 7);
    out.put(34, 24);
    out.put(kBlock, 16);
    out.put(kBlock, 16);
    out.put(0, 24);
    out.put(0, 24);
    out.put(uint64_t(rate), 20);
    out.put(uint64_t(channels - 1), 3);
    out.put(15, 5);
    out.put(uint64_t(frames), 36);
    // MD5 of the audio as encoded: `frames` frames of little-endian int16,
    // channels interleaved, zeros past the data frames.
    {
        Md5 md5;
        uint8_t sample[2] = {0, 0};
        for (int f = 0; f < frames; ++f)
            for (int c = 0; c < channels; ++c) {
                const int16_t v = f < dataFrames ? pcm[size_t(f) * size_t(channels) + size_t(c)] : int16_t(0);
                sample[0] = uint8_t(uint16_t(v) & 0xffu);
                sample[1] = uint8_t(uint16_t(v) >> 8);
                md5.update(sample, 2);
            }
        uint8_t digest[16];
        md5.finish(digest);
        for (int i = 0; i < 16; ++i)
            out.put(digest[i], 8);
    }
    std::vector<int64_t> ch[2], side, mid, best[4];
    uint64_t frameNumber = 0;
    for (int first = 0; first < frames; first += kBlock, ++frameNumber) {
        const int n = std::min(kBlock, frames - first);
        for (int c = 0; c < channels; ++c) {
            ch[c].assign(size_t(n), 0);
            for (int i = 0; i < n; ++i)
                if (first + i < dataFrames)
                    ch[c][size_t(i)] = pcm[size_t(first + i) * size_t(channels) + size_t(c)];
        }
        int assignment = channels - 1;
        std::vector<const std::vector<int64_t>*> signals;
        std::vector<int> depth;
        Subframe sub[2];
        std::vector<int64_t> res[2];
        if (channels == 1) {
            signals = {&ch[0]};
            depth = {16};
        } else {
            Subframe l = chooseSubframe(ch[0], 16, maxPartition, best[0]);
            Subframe r = chooseSubframe(ch[1], 16, maxPartition, best[1]);
            signals = {&ch[0], &ch[1]};
            depth = {16, 16};
            sub[0] = l;
            sub[1] = r;
            res[0] = best[0];
            res[1] = best[1];
            uint64_t bestBits = l.bits + r.bits;
            if (level >= 1) {
                side.resize(size_t(n));
                mid.resize(size_t(n));
                for (int i = 0; i < n; ++i) {
                    side[size_t(i)] = ch[0][size_t(i)] - ch[1][size_t(i)];
                    mid[size_t(i)] = (ch[0][size_t(i)] + ch[1][size_t(i)]) >> 1;
                }
                const Subframe s = chooseSubframe(side, 17, maxPartition, best[2]);
                const Subframe m = chooseSubframe(mid, 16, maxPartition, best[3]);
                if (l.bits + s.bits < bestBits) {
                    bestBits = l.bits + s.bits;
                    assignment = 8;
                    signals = {&ch[0], &side};
                    depth = {16, 17};
                    sub[0] = l, sub[1] = s, res[0] = best[0], res[1] = best[2];
                }
                if (s.bits + r.bits < bestBits) {
                    bestBits = s.bits + r.bits;
                    assignment = 9;
                    signals = {&side, &ch[1]};
                    depth = {17, 16};
                    sub[0] = s, sub[1] = r, res[0] = best[2], res[1] = best[1];
                }
                if (m.bits + s.bits < bestBits) {
                    assignment = 10;
                    signals = {&mid, &side};
                    depth = {16, 17};
                    sub[0] = m, sub[1] = s, res[0] = best[3], res[1] = best[2];
                }
            }
        }
        if (channels == 1) {
            sub[0] = chooseSubframe(ch[0], 16, maxPartition, best[0]);
            res[0] = best[0];
        }
        BitWriter f;
        f.put(0x3ffe, 14);
        f.put(0, 1);
        f.put(0, 1);
        f.put(7, 4);                    // 16-bit block size follows
        f.put(0, 4);                    // rate from STREAMINFO
        f.put(uint64_t(assignment), 4);
        f.put(4, 3);                    // 16 bits per sample
        f.put(0, 1);
        putUtf8(f, frameNumber);
        f.put(uint64_t(n - 1), 16);
        f.put(crc8(f.bytes.data(), f.bytes.size()), 8);
        for (size_t c = 0; c < signals.size(); ++c)
            writeSubframe(f, *signals[c], depth[c], sub[c], res[c]);
        f.align();
        const uint16_t crc = crc16(f.bytes.data(), f.bytes.size());
        f.put(crc, 16);
        out.bytes.insert(out.bytes.end(), f.bytes.begin(), f.bytes.end());
    }
    std::ofstream file(path, std::ios::binary);
    file.write(reinterpret_cast<const char*>(out.bytes.data()), std::streamsize(out.bytes.size()));
    file.close();
    if (!file)
        throw std::runtime_error("cannot write FLAC " + path);
}

bool readFlac(const std::string& path, int& rate, int& channels, std::vector<int16_t>& pcm)
{
    try {
        std::ifstream in(path, std::ios::binary);
        if (!in)
            return false;
        const std::string data{std::istreambuf_iterator<char>(in), {}};
        const auto* p = reinterpret_cast<const uint8_t*>(data.data());
        if (data.size() < 42 || std::memcmp(p, "fLaC", 4))
            return false;
        size_t at = 4;
        uint64_t total = 0;
        int bps = 0;
        for (bool last = false; !last;) {
            if (at + 4 > data.size())
                return false;
            last = p[at] & 0x80;
            const int type = p[at] & 0x7f;
            const size_t length = (size_t(p[at + 1]) << 16) | (size_t(p[at + 2]) << 8) | p[at + 3];
            if (at + 4 + length > data.size())
                return false;
            if (type == 0) {
                BitReader br(p + at + 4, length);
                br.get(16), br.get(16), br.get(24), br.get(24);
                rate = int(br.get(20));
                channels = int(br.get(3)) + 1;
                bps = int(br.get(5)) + 1;
                total = br.get(36);
            }
            at += 4 + length;
        }
        if (bps != 16 || !total)
            return false;
        pcm.assign(size_t(total) * size_t(channels), 0);
        uint64_t done = 0;
        std::vector<int64_t> x[2];
        while (done < total) {
            BitReader br(p + at, data.size() - at);
            if (br.get(14) != 0x3ffe)
                return false;
            br.get(1), br.get(1);
            const int sizeCode = int(br.get(4)), rateCode = int(br.get(4)), assignment = int(br.get(4));
            br.get(3), br.get(1);
            uint64_t lead = br.get(8);
            int extra = 0;
            while (lead & (0x80 >> extra))
                ++extra;
            for (int i = 1; i < extra; ++i)
                br.get(8);
            size_t n;
            if (sizeCode == 1) n = 192;
            else if (sizeCode >= 2 && sizeCode <= 5) n = size_t(576) << (sizeCode - 2);
            else if (sizeCode == 6) n = size_t(br.get(8)) + 1;
            else if (sizeCode == 7) n = size_t(br.get(16)) + 1;
            else if (sizeCode >= 8) n = size_t(256) << (sizeCode - 8);
            else return false;
            if (rateCode == 12) br.get(8);
            else if (rateCode == 13 || rateCode == 14) br.get(16);
            const size_t headerBytes = br.byte();
            if (crc8(p + at, headerBytes) != br.get(8))
                return false;
            const int nch = assignment < 8 ? assignment + 1 : 2;
            if (nch != channels || done + n > total)
                return false;
            for (int c = 0; c < nch; ++c) {
                const int depth = (assignment == 8 && c == 1) || (assignment == 9 && c == 0) || (assignment == 10 && c == 1) ? 17 : 16;
                readSubframe(br, n, depth, x[c]);
            }
            br.align();
            const size_t frameBytes = br.byte();
            if (crc16(p + at, frameBytes) != br.get(16))
                return false;
            for (size_t i = 0; i < n; ++i) {
                int64_t l = x[0][i], r = nch == 2 ? x[1][i] : 0;
                if (assignment == 8) r = l - r;
                else if (assignment == 9) l = l + r;
                else if (assignment == 10) {
                    const int64_t m = (l << 1) | (r & 1);
                    l = (m + r) >> 1;
                    r = (m - x[1][i]) >> 1;
                }
                pcm[(done + i) * size_t(channels)] = int16_t(l);
                if (nch == 2)
                    pcm[(done + i) * size_t(channels) + 1] = int16_t(r);
            }
            done += n;
            at += frameBytes + 2;
        }
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

} // namespace clay
