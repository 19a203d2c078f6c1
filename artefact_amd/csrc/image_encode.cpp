/*
 * RGB image -> file bytes: the last stage of the reference's CLI and WASM front ends, which
 * save the RgbImage that Artefact::process returns (backend/artefact-cli/main.rs:156,
 * img.save(output); backend/artefact-wasm/lib.rs:41, write_to(.., ImageFormat)).  The
 * reference does this through the `image` crate 0.25 with its png, bmp, tiff and webp
 * features (backend/artefact-lib/Cargo.toml:7-12), a dependency that is not vendored; the
 * formats here are written from their specifications:
 *
 *   PNG  (ISO/IEC 15948): 8-bit RGB (colour type 2), no interlace; every scanline takes the
 *        filter (None, Sub, Up, Average, Paeth) whose filtered bytes have the smallest sum of
 *        absolute values as signed bytes (the standard's adaptive heuristic), and the filtered
 *        data is one zlib stream (deflate at zlib's default level) split over IDAT chunks.
 *   BMP  (Windows bitmap, BITMAPINFOHEADER): 24 bits per pixel, BGR, bottom-up rows padded to
 *        4 bytes, uncompressed.
 *   TIFF (TIFF 6.0 baseline RGB, little-endian): one uncompressed strip, chunky RGB, 8 bits
 *        per sample, 72 dpi.
 *   WebP (lossless, VP8L bitstream of RFC 9649, in a RIFF container): the subtract-green
 *        transform, no colour cache, no backward references; one prefix code per channel
 *        (Huffman code lengths from the image's histograms, limited to 15 bits, written as
 *        a "simple" code for one or two symbols, else a "normal" code whose code-length code
 *        uses zero runs).  The image crate's WebP encoder is lossless too.
 *
 * Host C++; nothing here touches the GPU.
 */
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/artefact_hip.h"

namespace {

constexpr size_t PNG_CHUNK = 1u << 20;  // IDAT payload per chunk

struct Writer {
    uint8_t *out;
    size_t cap, n = 0;
    bool ok = true;
    void bytes(const void *p, size_t k)
    {
        if (!ok || k > cap - n) {
            ok = false;
            return;
        }
        memcpy(out + n, p, k);
        n += k;
    }
    void u8(uint8_t v) { bytes(&v, 1); }
    void be32(uint32_t v)
    {
        const uint8_t b[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
        bytes(b, 4);
    }
    void le16(uint16_t v)
    {
        const uint8_t b[2] = {(uint8_t)v, (uint8_t)(v >> 8)};
        bytes(b, 2);
    }
    void le32(uint32_t v)
    {
        const uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        bytes(b, 4);
    }
};

// One PNG chunk: length, type, data, CRC-32 of type and data.
void png_chunk(Writer &w, const char type[4], const uint8_t *data, size_t len)
{
    w.be32((uint32_t)len);
    w.bytes(type, 4);
    if (len) w.bytes(data, len);
    uLong crc = crc32(0L, reinterpret_cast<const Bytef *>(type), 4);
    if (len) crc = crc32(crc, data, (uInt)len);
    w.be32((uint32_t)crc);
}

inline int paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Filters one scanline (bpp = 3) with filter type t into dst (length n); prev is the previous
// unfiltered scanline (zeros for the first).
void png_filter(int t, const uint8_t *cur, const uint8_t *prev, size_t n, uint8_t *dst)
{
    for (size_t i = 0; i < n; i++) {
        const int a = i >= 3 ? cur[i - 3] : 0, b = prev[i], c = i >= 3 ? prev[i - 3] : 0;
        int pred = 0;
        switch (t) {
        case 1: pred = a; break;
        case 2: pred = b; break;
        case 3: pred = (a + b) >> 1; break;
        case 4: pred = paeth(a, b, c); break;
        default: break;
        }
        dst[i] = (uint8_t)(cur[i] - pred);
    }
}

size_t png_bound(uint32_t w, uint32_t h)
{
    const uint64_t raw = (uint64_t)h * (1 + 3ull * w);
    const uint64_t z = compressBound((uLong)raw);
    return (size_t)(8 + 25 + 12 + z + 12 * (z / PNG_CHUNK + 1));
}

int encode_png(const uint8_t *rgb, uint32_t w, uint32_t h, Writer &out)
{
    const size_t stride = 3ull * w;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    out.bytes(sig, 8);
    uint8_t ihdr[13];
    const uint32_t be[2] = {w, h};
    for (int k = 0; k < 2; k++)
        for (int b = 0; b < 4; b++) ihdr[4 * k + b] = (uint8_t)(be[k] >> (24 - 8 * b));
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 2;   // colour type: RGB
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering with five basic types
    ihdr[12] = 0;  // no interlace
    png_chunk(out, "IHDR", ihdr, 13);

    std::vector<uint8_t> zero(stride, 0), line(stride + 1), trial(stride);
    std::vector<uint8_t> zbuf(PNG_CHUNK);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (deflateInit(&zs, Z_DEFAULT_COMPRESSION) != Z_OK) return ARTEFACT_E_HIP;
    int rc = ARTEFACT_OK;
    // Runs deflate over the pending input (mode Z_NO_FLUSH) or to the end of the stream
    // (Z_FINISH), writing an IDAT chunk whenever the output buffer fills and, at the end, one
    // for what is left.
    auto pump = [&](int mode) {
        for (;;) {
            const int r = deflate(&zs, mode);
            if (r == Z_STREAM_ERROR) return false;
            const bool done = mode == Z_FINISH ? r == Z_STREAM_END : zs.avail_in == 0 && zs.avail_out != 0;
            const size_t have = PNG_CHUNK - zs.avail_out;
            if (zs.avail_out == 0 || (done && mode == Z_FINISH && have)) {
                png_chunk(out, "IDAT", zbuf.data(), have);
                zs.next_out = zbuf.data();
                zs.avail_out = (uInt)PNG_CHUNK;
            }
            if (done) return true;
        }
    };
    zs.next_out = zbuf.data();
    zs.avail_out = (uInt)PNG_CHUNK;
    for (uint32_t y = 0; y < h && rc == ARTEFACT_OK; y++) {
        const uint8_t *cur = rgb + (size_t)y * stride;
        const uint8_t *prev = y ? cur - stride : zero.data();
        uint64_t best = UINT64_MAX;
        for (int t = 0; t < 5; t++) {
            png_filter(t, cur, prev, stride, trial.data());
            uint64_t s = 0;
            for (size_t i = 0; i < stride; i++) s += (uint64_t)abs((int)(int8_t)trial[i]);
            if (s < best) {
                best = s;
                line[0] = (uint8_t)t;
                memcpy(line.data() + 1, trial.data(), stride);
            }
        }
        zs.next_in = line.data();
        zs.avail_in = (uInt)(stride + 1);
        if (!pump(Z_NO_FLUSH)) rc = ARTEFACT_E_HIP;
    }
    if (rc == ARTEFACT_OK && !pump(Z_FINISH)) rc = ARTEFACT_E_HIP;
    deflateEnd(&zs);
    if (rc) return rc;
    png_chunk(out, "IEND", nullptr, 0);
    return out.ok ? ARTEFACT_OK : ARTEFACT_E_ARGUMENT;
}

size_t bmp_row(uint32_t w) { return (3ull * w + 3) & ~(size_t)3; }

int encode_bmp(const uint8_t *rgb, uint32_t w, uint32_t h, Writer &out)
{
    const size_t row = bmp_row(w), img = row * h, off = 14 + 40;
    if (off + img > 0xFFFFFFFFull) return ARTEFACT_E_ARGUMENT;
    out.u8('B');
    out.u8('M');
    out.le32((uint32_t)(off + img));
    out.le32(0);
    out.le32((uint32_t)off);
    out.le32(40);           // BITMAPINFOHEADER
    out.le32(w);
    out.le32(h);            // positive: bottom-up rows
    out.le16(1);            // planes
    out.le16(24);           // bits per pixel
    out.le32(0);            // BI_RGB
    out.le32((uint32_t)img);
    out.le32(2835);         // 72 dpi
    out.le32(2835);
    out.le32(0);
    out.le32(0);
    std::vector<uint8_t> line(row, 0);
    for (uint32_t y = h; y-- > 0;) {
        const uint8_t *src = rgb + (size_t)y * 3 * w;
        for (uint32_t x = 0; x < w; x++) {
            line[3 * x + 0] = src[3 * x + 2];
            line[3 * x + 1] = src[3 * x + 1];
            line[3 * x + 2] = src[3 * x + 0];
        }
        out.bytes(line.data(), row);
    }
    return out.ok ? ARTEFACT_OK : ARTEFACT_E_ARGUMENT;
}

constexpr uint32_t TIFF_TAGS = 12;
constexpr size_t TIFF_HEADER = 8 + 2 + 12 * TIFF_TAGS + 4 + 6 + 16;  // header, IFD, BitsPerSample, 2 rationals

int encode_tiff(const uint8_t *rgb, uint32_t w, uint32_t h, Writer &out)
{
    const uint64_t img = 3ull * w * h;
    if (TIFF_HEADER + img > 0xFFFFFFFFull) return ARTEFACT_E_ARGUMENT;
    const uint32_t ifd = 8, bps = ifd + 2 + 12 * TIFF_TAGS + 4, res = bps + 6, data = res + 16;
    out.u8('I');
    out.u8('I');
    out.le16(42);
    out.le32(ifd);
    out.le16(TIFF_TAGS);
    auto tag = [&](uint16_t id, uint16_t type, uint32_t count, uint32_t value) {
        out.le16(id);
        out.le16(type);
        out.le32(count);
        if (type == 3 && count == 1) {  // a SHORT sits in the low half of the value field
            out.le16((uint16_t)value);
            out.le16(0);
        } else {
            out.le32(value);
        }
    };
    // Entries in increasing tag order (TIFF 6.0 section 2).
    tag(256, 4, 1, w);                    // ImageWidth (LONG)
    tag(257, 4, 1, h);                    // ImageLength
    tag(258, 3, 3, bps);                  // BitsPerSample: 8, 8, 8 at bps
    tag(259, 3, 1, 1);                    // Compression: none
    tag(262, 3, 1, 2);                    // PhotometricInterpretation: RGB
    tag(273, 4, 1, data);                 // StripOffsets
    tag(277, 3, 1, 3);                    // SamplesPerPixel
    tag(278, 4, 1, h);                    // RowsPerStrip: one strip
    tag(279, 4, 1, (uint32_t)img);        // StripByteCounts
    tag(282, 5, 1, res);                  // XResolution (RATIONAL)
    tag(283, 5, 1, res + 8);              // YResolution
    tag(296, 3, 1, 2);                    // ResolutionUnit: inch
    out.le32(0);                          // no next IFD
    out.le16(8);
    out.le16(8);
    out.le16(8);
    out.le32(72);
    out.le32(1);
    out.le32(72);
    out.le32(1);
    out.bytes(rgb, (size_t)img);
    return out.ok ? ARTEFACT_OK : ARTEFACT_E_ARGUMENT;
}


// ---- WebP lossless (VP8L) ----

// Bits in the order VP8L reads them: least significant first.
struct BitWriter {
    std::vector<uint8_t> buf;
    uint64_t acc = 0;
    int nbits = 0;
    void put(uint32_t v, int n)
    {
        acc |= (uint64_t)v << nbits;
        nbits += n;
        while (nbits >= 8) {
            buf.push_back((uint8_t)acc);
            acc >>= 8;
            nbits -= 8;
        }
    }
    void flush()
    {
        if (nbits > 0) buf.push_back((uint8_t)acc);
        acc = 0;
        nbits = 0;
    }
};

// Huffman code lengths for the counts, none longer than max_len; unused symbols get 0.  Ties
// break on the node index, so the lengths are deterministic.  A tree deeper than max_len is
// rebuilt from halved counts until it fits (equal counts give a balanced tree).
std::vector<uint8_t> huffman_lengths(const std::vector<uint64_t> &count, int max_len)
{
    const size_t n = count.size();
    std::vector<uint8_t> len(n, 0);
    std::vector<uint64_t> c(count);
    for (;;) {
        std::vector<std::pair<int, int>> kids;  // internal node k: children (node ids)
        std::vector<int> leaf_symbol;           // node id < leaves: the symbol
        using Item = std::pair<uint64_t, int>;  // (weight, node id)
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
        for (size_t s = 0; s < n; s++)
            if (c[s]) {
                heap.push({c[s], (int)leaf_symbol.size()});
                leaf_symbol.push_back((int)s);
            }
        const int leaves = (int)leaf_symbol.size();
        if (leaves == 0) return len;
        if (leaves == 1) {
            len[leaf_symbol[0]] = 1;
            return len;
        }
        while (heap.size() > 1) {
            const Item a = heap.top();
            heap.pop();
            const Item b = heap.top();
            heap.pop();
            kids.push_back({a.second, b.second});
            heap.push({a.first + b.first, leaves + (int)kids.size() - 1});
        }
        // Depths from the root down.
        std::vector<int> depth(leaves + kids.size(), 0);
        int deepest = 0;
        for (int k = (int)kids.size() - 1; k >= 0; k--) {
            const int d = depth[leaves + k] + 1;
            depth[kids[k].first] = d;
            depth[kids[k].second] = d;
        }
        for (int i = 0; i < leaves; i++) deepest = std::max(deepest, depth[i]);
        if (deepest <= max_len) {
            for (int i = 0; i < leaves; i++) len[leaf_symbol[i]] = (uint8_t)depth[i];
            return len;
        }
        for (uint64_t &v : c)
            if (v) v = (v >> 1) | 1;
    }
}

// Canonical codes for the lengths (shorter codes first, then by symbol), each bit-reversed so
// that writing it least significant bit first puts its first bit first.
std::vector<uint32_t> canonical_codes(const std::vector<uint8_t> &len)
{
    uint32_t count[16] = {}, next[16] = {};
    for (uint8_t l : len)
        if (l) count[l]++;
    uint32_t code = 0;
    for (int l = 1; l < 16; l++) {
        code = (code + count[l - 1]) << 1;
        next[l] = code;
    }
    std::vector<uint32_t> out(len.size(), 0);
    for (size_t s = 0; s < len.size(); s++) {
        const int l = len[s];
        if (!l) continue;
        const uint32_t c = next[l]++;
        uint32_t r = 0;
        for (int i = 0; i < l; i++) r |= ((c >> i) & 1u) << (l - 1 - i);
        out[s] = r;
    }
    return out;
}

struct PrefixCode {
    std::vector<uint8_t> len;
    std::vector<uint32_t> code;
    void emit(BitWriter &bw, int s) const { bw.put(code[s], len[s]); }
};

// Writes the prefix code for an alphabet with these symbol counts and returns it.
PrefixCode write_prefix_code(BitWriter &bw, const std::vector<uint64_t> &count)
{
    const size_t n = count.size();
    PrefixCode p;
    p.len.assign(n, 0);
    p.code.assign(n, 0);
    std::vector<int> used;
    for (size_t s = 0; s < n; s++)
        if (count[s]) used.push_back((int)s);
    if (used.size() <= 2 && (used.empty() || used.back() < 256)) {
        // Simple code: one symbol (read with no bits) or two (one bit each, the smaller first).
        if (used.empty()) used.push_back(0);
        bw.put(1, 1);
        bw.put((uint32_t)used.size() - 1, 1);
        if (used[0] < 2) {
            bw.put(0, 1);
            bw.put((uint32_t)used[0], 1);
        } else {
            bw.put(1, 1);
            bw.put((uint32_t)used[0], 8);
        }
        if (used.size() == 2) {
            bw.put((uint32_t)used[1], 8);
            p.len[used[0]] = p.len[used[1]] = 1;
            p.code[used[1]] = 1;
        }
        return p;
    }
    p.len = huffman_lengths(count, 15);
    p.code = canonical_codes(p.len);
    // The code lengths as tokens: a length 0-15, or a run of zeros (17: 3-10, 18: 11-138).
    std::vector<std::pair<int, int>> tokens;  // (symbol, extra bits value)
    for (size_t i = 0; i < n;) {
        if (p.len[i]) {
            tokens.push_back({p.len[i], 0});
            i++;
            continue;
        }
        size_t run = 1;
        while (i + run < n && !p.len[i + run]) run++;
        i += run;
        while (run >= 11) {
            const size_t k = std::min<size_t>(run, 138);
            tokens.push_back({18, (int)(k - 11)});
            run -= k;
        }
        if (run >= 3) {
            tokens.push_back({17, (int)(run - 3)});
            run = 0;
        }
        for (; run; run--) tokens.push_back({0, 0});
    }
    std::vector<uint64_t> hist(19, 0);
    for (const auto &t : tokens) hist[t.first]++;
    // A code-length code of one symbol would be a one-symbol normal code; a second symbol that
    // is never written keeps it a complete two-symbol code.
    int distinct = 0;
    for (uint64_t v : hist) distinct += v ? 1 : 0;
    if (distinct == 1) hist[hist[0] ? 1 : 0] = 1;
    const std::vector<uint8_t> cl_len = huffman_lengths(hist, 7);
    const std::vector<uint32_t> cl_code = canonical_codes(cl_len);
    static const int order[19] = {17, 18, 0, 1, 2, 3, 4, 5, 16, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    int num = 19;
    while (num > 4 && !cl_len[order[num - 1]]) num--;
    bw.put(0, 1);  // normal code
    bw.put((uint32_t)(num - 4), 4);
    for (int i = 0; i < num; i++) bw.put(cl_len[order[i]], 3);
    bw.put(0, 1);  // the code lengths of the whole alphabet follow
    for (const auto &t : tokens) {
        bw.put(cl_code[t.first], cl_len[t.first]);
        if (t.first == 17) bw.put((uint32_t)t.second, 3);
        else if (t.first == 18) bw.put((uint32_t)t.second, 7);
    }
    return p;
}

constexpr uint32_t WEBP_MAX_SIDE = 16384;  // 14-bit width - 1 and height - 1

size_t webp_bound(uint32_t w, uint32_t h)
{
    // At most 15 bits for each of green, red and blue per pixel (alpha is one symbol: no bits),
    // plus headers and the five code descriptions.
    return (size_t)w * h * 6 + 8192;
}

int encode_webp(const uint8_t *rgb, uint32_t w, uint32_t h, Writer &out)
{
    if (w > WEBP_MAX_SIDE || h > WEBP_MAX_SIDE) return ARTEFACT_E_ARGUMENT;
    const size_t npx = (size_t)w * h;
    // Subtract-green transform: red and blue minus green, modulo 256.
    std::vector<uint64_t> hg(256 + 24, 0), hr(256, 0), hb(256, 0), ha(256, 0), hd(40, 0);
    for (size_t i = 0; i < npx; i++) {
        const uint8_t r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        hg[g]++;
        hr[(uint8_t)(r - g)]++;
        hb[(uint8_t)(b - g)]++;
    }
    ha[255] = npx;
    BitWriter bw;
    bw.put(0x2f, 8);  // VP8L signature
    bw.put(w - 1, 14);
    bw.put(h - 1, 14);
    bw.put(0, 1);  // alpha_is_used: every alpha is 255
    bw.put(0, 3);  // version 0
    bw.put(1, 1);  // a transform follows:
    bw.put(2, 2);  // subtract green
    bw.put(0, 1);  // no further transform
    bw.put(0, 1);  // no colour cache
    bw.put(0, 1);  // no meta prefix codes
    const PrefixCode cg = write_prefix_code(bw, hg), cr = write_prefix_code(bw, hr), cb = write_prefix_code(bw, hb),
                     ca = write_prefix_code(bw, ha);
    (void)write_prefix_code(bw, hd);  // distances: no backward references are written
    for (size_t i = 0; i < npx; i++) {
        const uint8_t r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        cg.emit(bw, g);
        cr.emit(bw, (uint8_t)(r - g));
        cb.emit(bw, (uint8_t)(b - g));
        ca.emit(bw, 255);
    }
    bw.flush();
    const size_t data = bw.buf.size(), padded = data + (data & 1);
    if (12 + 8 + padded > 0xFFFFFFFFull) return ARTEFACT_E_ARGUMENT;
    out.bytes("RIFF", 4);
    out.le32((uint32_t)(4 + 8 + padded));
    out.bytes("WEBP", 4);
    out.bytes("VP8L", 4);
    out.le32((uint32_t)data);
    out.bytes(bw.buf.data(), data);
    if (data & 1) out.u8(0);
    return out.ok ? ARTEFACT_OK : ARTEFACT_E_ARGUMENT;
}

}  // namespace

extern "C" {

size_t artefact_image_encode_bound(uint32_t w, uint32_t h, uint32_t format)
{
    if (!w || !h) return 0;
    switch (format) {
    case ARTEFACT_IMAGE_PNG: return png_bound(w, h);
    case ARTEFACT_IMAGE_BMP: return 14 + 40 + bmp_row(w) * h;
    case ARTEFACT_IMAGE_TIFF: return TIFF_HEADER + 3ull * w * h;
    case ARTEFACT_IMAGE_WEBP: return w > WEBP_MAX_SIDE || h > WEBP_MAX_SIDE ? 0 : webp_bound(w, h);
    default: return 0;
    }
}

int artefact_image_encode(const uint8_t *rgb, uint32_t w, uint32_t h, uint32_t format, uint8_t *out, size_t cap,
                          size_t *out_len)
{
    if (!rgb || !out || !out_len || !w || !h) return ARTEFACT_E_ARGUMENT;
    *out_len = 0;
    Writer wr{out, cap};
    int rc;
    try {
        switch (format) {
        case ARTEFACT_IMAGE_PNG: rc = encode_png(rgb, w, h, wr); break;
        case ARTEFACT_IMAGE_BMP: rc = encode_bmp(rgb, w, h, wr); break;
        case ARTEFACT_IMAGE_TIFF: rc = encode_tiff(rgb, w, h, wr); break;
        case ARTEFACT_IMAGE_WEBP: rc = encode_webp(rgb, w, h, wr); break;
        default: return ARTEFACT_E_ARGUMENT;
        }
    } catch (const std::bad_alloc &) {
        return ARTEFACT_E_HIP;
    }
    if (rc == ARTEFACT_OK) *out_len = wr.n;
    return rc;
}

}  // extern "C"
