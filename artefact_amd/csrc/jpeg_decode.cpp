/*
 * JPEG -> quantised DCT coefficients: the input stage of Artefact::process
 * (backend/artefact-lib/jpeg/zune.rs:5-62 over the zune-jpeg fork in backend/zune-jpeg).
 *
 * Host C++ (serial bitstream work stays on the CPU).  Supports what the fork supports:
 * sequential Huffman (SOF0, SOF1) and progressive Huffman (SOF2), 8-bit samples, 1 to 4
 * components with sampling factors 1 or 2 (zune-jpeg/src/components.rs:80-112), quantisation
 * tables of 8 or 16 bits (headers.rs:115-179), restart intervals.  The output follows the
 * fork's layout, not only the standard's:
 *
 *   - Every component plane is the image rounded up to the MCU, divided by the component's
 *     upsampling factor max/own (misc.rs:206-242); coefficients are stored in raster block
 *     order over that plane, natural (row-major) order inside a block (bitstream.rs:363,378),
 *     not dequantised.  Quantisation tables come out in natural order (headers.rs:540-552).
 *   - Sequential files: only the first scan is decoded, and every MCU holds every component
 *     (mcu.rs:80-285).  A component sampled (h=1, v=2) has its two blocks of an MCU stored
 *     at 2*mcu and 2*mcu+1 (mcu.rs:235-249) rather than one block row apart; blocks whose
 *     slot lies past the plane are skipped without reading bits (bitstream.rs:338-340).  A
 *     premature end (more than 37 bytes read past the data) stops decoding and keeps the
 *     zeros (mcu.rs:206-218).
 *   - DC predictions accumulate in 32 bits and are stored truncated to 16 bits
 *     (bitstream.rs:300-302, 430-432); progressive DC and AC values are shifted by Al in
 *     16 bits.
 *   - A single-component image whose component is subsampled keeps the plane size of its
 *     sampling factors but is decoded as 1x1 blocks (mcu_prog.rs:415-433).
 *
 * Everything else is ITU-T T.81: Annex B syntax, Annex C Huffman tables, F.2.2 sequential
 * decoding, G.1.2 progressive decoding (spectral selection and successive approximation).
 */
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/artefact_hip.h"

namespace {

// Zigzag scan position -> natural row-major index (T.81 Figure A.6).
constexpr uint8_t kNatural[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                  35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                  58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Error {
    std::string msg;
};

// Largest accepted width and height, in pixels.
constexpr int MAX_SIDE = 1 << 14;

struct Huffman {
    bool set = false;
    int maxcode[18];  // largest code of each length, -1 if none (F.2.2.3)
    int valptr[17];
    int mincode[17];
    uint8_t vals[256];
    uint16_t fast[1 << 9];  // 9-bit lookahead: (length << 8) | symbol, 0 if longer
};

void build_huffman(Huffman &h, const uint8_t counts[16], const uint8_t *vals, int nvals)
{
    memcpy(h.vals, vals, (size_t)nvals);
    memset(h.fast, 0, sizeof h.fast);
    int code = 0, k = 0;
    for (int l = 1; l <= 16; l++) {
        h.valptr[l] = k;
        h.mincode[l] = code;
        const int n = counts[l - 1];
        if (n == 0) {
            h.maxcode[l] = -1;
        } else {
            if (code + n > (1 << l)) throw Error{"Huffman table has more codes than its lengths allow"};
            for (int i = 0; i < n; i++, code++, k++) {
                if (l <= 9) {
                    const int lo = code << (9 - l), hi = (code + 1) << (9 - l);
                    for (int x = lo; x < hi; x++) h.fast[x] = (uint16_t)((l << 8) | vals[k]);
                }
            }
            h.maxcode[l] = code - 1;
        }
        code <<= 1;
    }
    h.maxcode[17] = 0x7fffffff;
    h.set = true;
}

// Entropy-coded segment reader: removes stuffed zero bytes, stops at a marker (then feeds
// zero bits and remembers the marker), counts the bytes it feeds past the end of the data.
struct BitReader {
    const uint8_t *d = nullptr;
    size_t n = 0, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    int marker = -1;
    size_t overread = 0;

    void fill()
    {
        while (cnt <= 56) {
            uint32_t b = 0;
            if (marker < 0 && pos < n) {
                b = d[pos++];
                if (b == 0xFF) {
                    size_t q = pos;
                    while (q < n && d[q] == 0xFF) q++;
                    if (q >= n) {
                        marker = 0xD9;
                        pos = n;
                        b = 0;
                    } else if (d[q] == 0x00) {
                        pos = q + 1;
                    } else {
                        marker = d[q];
                        pos = q + 1;
                        b = 0;
                    }
                }
            } else if (pos >= n) {
                overread++;  // past the end of the data (zeros after a marker do not count)
            }
            buf |= (uint64_t)b << (56 - cnt);
            cnt += 8;
        }
    }
    uint32_t peek(int k)
    {
        if (cnt < k) fill();
        return (uint32_t)(buf >> (64 - k));
    }
    void drop(int k)
    {
        buf <<= k;
        cnt -= k;
    }
    uint32_t get(int k)
    {
        if (k == 0) return 0;
        const uint32_t v = peek(k);
        drop(k);
        return v;
    }
    void reset_bits()
    {
        buf = 0;
        cnt = 0;
    }
};

int extend(uint32_t v, int s) { return (v < (1u << (s - 1))) ? (int)v - (1 << s) + 1 : (int)v; }

// Decodes one Huffman symbol.  If `lookup` is given it receives whether the fork's combined
// run/value table would have decoded an AC symbol (huffman.rs:189-238): a code of at most 9
// bits whose symbol index is below 255, with magnitude bits s != 0, code length + s <= 9 and an
// extended value in [-128, 127].  The caller checks the value; here the first three conditions
// are stored as the symbol's fast-table length (0 when not eligible).
int decode_symbol(BitReader &br, const Huffman &h, int *fast_len = nullptr)
{
    const uint32_t look = br.peek(16);
    const uint16_t f = h.fast[look >> 7];
    if (f) {
        const int l = f >> 8;
        br.drop(l);
        if (fast_len) {
            const int code = (int)(look >> (16 - l));
            const int index = h.valptr[l] + code - h.mincode[l];
            *fast_len = index < 255 ? l : 0;
        }
        return f & 0xFF;
    }
    for (int l = 10; l <= 16; l++) {
        const int code = (int)(look >> (16 - l));
        if (code <= h.maxcode[l]) {
            br.drop(l);
            if (fast_len) *fast_len = 0;
            return h.vals[h.valptr[l] + code - h.mincode[l]];
        }
    }
    throw Error{"bad Huffman code"};
}

// Natural index at which the fork stores an AC value decoded at zigzag position k, which can
// pass 63 in a corrupt stream.  Its combined table path clamps (UN_ZIGZAG[min(k, 63)]), its
// other path wraps (UN_ZIGZAG[k & 63]): bitstream.rs:359-378 (sequential) and :487-503
// (progressive first scans).
int ac_position(int k, int fast_len, int s, int value)
{
    const bool fast = fast_len != 0 && s != 0 && fast_len + s <= 9 && value >= -128 && value <= 127;
    return kNatural[fast ? std::min(k, 63) : (k & 63)];
}

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;  // tables of the current scan
    uint32_t rounded_w = 0, rounded_h = 0;
    std::vector<int16_t> coef;
    int32_t dc_pred = 0;
};

struct Decoder {
    const uint8_t *d;
    size_t n, pos = 0;
    uint16_t qt[4][64] = {};
    bool qt_set[4] = {false, false, false, false};
    Huffman dc[4], ac[4];
    int width = 0, height = 0;
    std::vector<Component> comp;
    int hmax = 1, vmax = 1;
    bool interleaved = false, progressive = false, seen_sof = false, grey_reset = false;
    int restart_interval = 0;
    // current scan
    std::vector<int> scomp;
    int ss = 0, se = 63, ah = 0, al = 0;
    int eobrun = 0;
    int mcu_w = 0, mcu_h = 0;  // min_mcu_w / min_mcu_h of misc.rs:217-220

    Decoder(const uint8_t *data, size_t size) : d(data), n(size) {}

    uint8_t byte()
    {
        if (pos >= n) throw Error{"unexpected end of data"};
        return d[pos++];
    }
    uint16_t u16()
    {
        const uint16_t hi = byte();
        return (uint16_t)((hi << 8) | byte());
    }
    int next_marker()
    {
        // Skips anything up to 0xFF xx with xx not 0x00 or 0xFF.
        while (pos < n) {
            if (d[pos++] != 0xFF) continue;
            while (pos < n && d[pos] == 0xFF) pos++;
            if (pos >= n) break;
            const uint8_t m = d[pos++];
            if (m != 0x00) return m;
        }
        return -1;
    }

    void parse_dqt()
    {
        int len = (int)u16() - 2;
        while (len > 0) {
            const uint8_t pq = byte();
            const int prec = pq >> 4, t = pq & 15;
            if (prec > 1) throw Error{"quantisation table precision must be 0 or 1"};
            if (t > 3) throw Error{"quantisation table id above 3"};
            for (int k = 0; k < 64; k++) qt[t][kNatural[k]] = prec ? u16() : byte();
            qt_set[t] = true;
            len -= 1 + 64 * (prec + 1);
        }
        if (len != 0) throw Error{"bad DQT length"};
    }
    void parse_dht()
    {
        int len = (int)u16() - 2;
        while (len > 0) {
            const uint8_t tc_th = byte();
            const int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3) throw Error{"bad Huffman table class or id"};
            uint8_t counts[16];
            int total = 0;
            for (int i = 0; i < 16; i++) total += counts[i] = byte();
            if (total > 256) throw Error{"Huffman table with more than 256 symbols"};
            uint8_t vals[256];
            for (int i = 0; i < total; i++) vals[i] = byte();
            build_huffman(tc ? ac[th] : dc[th], counts, vals, total);
            len -= 17 + total;
        }
        if (len != 0) throw Error{"bad DHT length"};
    }
    void parse_sof(int marker)
    {
        if (seen_sof) throw Error{"more than one frame header"};
        seen_sof = true;
        progressive = marker == 0xC2;
        const int len = u16();
        if (byte() != 8) throw Error{"only 8-bit samples are supported"};
        height = u16();
        width = u16();
        const int nc = byte();
        // The size limits of the decoder's default options, checked before the zero size
        // (zune-jpeg/src/headers.rs:216-230; zune.rs:12 builds the decoder with the defaults
        // of zune-core 0.5.0-rc2, 1 << 14 per side).
        if (width > MAX_SIDE)
            throw Error{"image width " + std::to_string(width) + " greater than width limit " +
                        std::to_string(MAX_SIDE)};
        if (height > MAX_SIDE)
            throw Error{"image height " + std::to_string(height) + " greater than height limit " +
                        std::to_string(MAX_SIDE)};
        if (width == 0 || height == 0) throw Error{"zero image size"};
        if (nc < 1 || nc > 4) throw Error{"component count must be 1 to 4"};
        if (len != 8 + 3 * nc) throw Error{"bad SOF length"};
        comp.assign(nc, Component{});
        for (int i = 0; i < nc; i++) {
            comp[i].id = byte();
            const uint8_t hv = byte();
            comp[i].h = hv >> 4;
            comp[i].v = hv & 15;
            comp[i].tq = byte();
            if (comp[i].h < 1 || comp[i].h > 2 || comp[i].v < 1 || comp[i].v > 2)
                throw Error{"sampling factors must be 1 or 2"};
            if (comp[i].tq > 3) throw Error{"quantisation table id above 3"};
        }
    }
    void parse_sos()
    {
        const int len = u16();
        const int ns = byte();
        if (ns < 1 || ns > (int)comp.size() || len != 6 + 2 * ns) throw Error{"bad SOS header"};
        scomp.clear();
        for (int i = 0; i < ns; i++) {
            const int cid = byte(), tables = byte();
            int k = -1;
            for (size_t j = 0; j < comp.size(); j++)
                if (comp[j].id == cid) k = (int)j;
            if (k < 0) throw Error{"scan names an unknown component"};
            comp[k].td = tables >> 4;
            comp[k].ta = tables & 15;
            if (comp[k].td > 3 || comp[k].ta > 3) throw Error{"Huffman table id above 3"};
            scomp.push_back(k);
        }
        ss = byte();
        se = byte();
        const int a = byte();
        ah = a >> 4;
        al = a & 15;
        if (progressive) {
            if (ss > 63 || se > 63 || ss > se || (ss == 0 && se != 0) || (ss != 0 && ns != 1) || al > 13)
                throw Error{"bad progressive scan parameters"};
        }
    }

    // Plane geometry and coefficient storage (misc.rs:206-242, decoder.rs:190-193).
    void setup()
    {
        hmax = vmax = 1;
        for (const Component &c : comp) {
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        interleaved = hmax != 1 || vmax != 1;
        const uint32_t mw = 8u * hmax - 1, mh = 8u * vmax - 1;
        const uint32_t rw = ((uint32_t)width + mw) & ~mw, rh = ((uint32_t)height + mh) & ~mh;
        mcu_w = (width + 8 * hmax - 1) / (8 * hmax);
        mcu_h = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component &c : comp) {
            if (!qt_set[c.tq]) throw Error{"no quantisation table for a component"};
            c.rounded_w = rw / (uint32_t)(hmax / c.h);
            c.rounded_h = rh / (uint32_t)(vmax / c.v);
            c.coef.assign((size_t)c.rounded_w * c.rounded_h, 0);
        }
        if (comp.size() == 1 && interleaved) {
            // A grey image with a subsampled component: decoded as 1x1 (mcu_prog.rs:415-433).
            grey_reset = true;
            comp[0].h = comp[0].v = 1;
            hmax = vmax = 1;
            interleaved = false;
        }
    }

    void decode_block_sequential(BitReader &br, Component &c, int16_t *blk)
    {
        const Huffman &hd = dc[c.td], &ha = ac[c.ta];
        if (!hd.set || !ha.set) throw Error{"missing Huffman table"};
        const int t = decode_symbol(br, hd);
        if (t > 16) throw Error{"bad DC magnitude"};
        const int diff = t ? extend(br.get(t), t) : 0;
        c.dc_pred = (int32_t)((uint32_t)c.dc_pred + (uint32_t)diff);
        blk[0] = (int16_t)c.dc_pred;
        for (int k = 1; k < 64;) {
            int fl = 0;
            const int rs = decode_symbol(br, ha, &fl);
            const int r = rs >> 4, s = rs & 15;
            if (s) {
                k += r;
                const int v = extend(br.get(s), s);
                blk[ac_position(k, fl, s, v)] = (int16_t)v;
                k++;
            } else if (r != 15) {
                break;
            } else {
                k += 16;
            }
        }
    }

    // Restart: the bit buffer, the pending marker, the end-of-band run and the DC predictions
    // start over (bitstream.rs:669-675, mcu.rs:339-360).
    void reset_at_marker(BitReader &br)
    {
        br.reset_bits();
        br.marker = -1;
        for (Component &c : comp) c.dc_pred = 0;
        eobrun = 0;
    }
    static bool is_rst(int m) { return m >= 0xD0 && m <= 0xD7; }

    // Sequential scans, after every MCU (mcu.rs:288-330): the countdown stops at 0, and the
    // restart happens at the end of the first MCU at which the countdown is 0 and the bit
    // reader has met an RSTn marker.  In a valid file that is the interval's last MCU; when the
    // marker is missing, decoding continues without a restart until the next one is met.  With
    // no restart interval the countdown is always 0, so a stray RSTn restarts the stream.
    void after_sequential_mcu(BitReader &br, int &todo)
    {
        if (todo > 0) todo--;
        if (todo == 0 && is_rst(br.marker)) {
            reset_at_marker(br);
            todo = restart_interval;
        }
    }

    // Progressive scans, after every MCU or block (mcu_prog.rs:323-328, 409-414): at the end of
    // every interval the countdown starts again, and the restart happens only if the reader has
    // met an RSTn marker by then (mcu.rs:339-360); a missing marker is not searched for.  The
    // countdown starts with each scan, as T.81 defines restart intervals (the fork carries it
    // over from the previous scan and rejects a header marker met at an interval's end; with
    // lazy refills, whether it has met one there depends on its bit buffer, which this decoder
    // does not model: a scan ends with its data here).
    void after_progressive_unit(BitReader &br, int &todo)
    {
        if (!restart_interval || --todo != 0) return;
        todo = restart_interval;
        if (is_rst(br.marker)) reset_at_marker(br);
    }

    void decode_sequential(BitReader &br)
    {
        const int mw = interleaved ? mcu_w : (width + 7) / 8;
        const int mh = interleaved ? mcu_h : (height + 7) / 8;
        int todo = restart_interval;
        for (int r = 0; r < mh; r++) {
            if (br.overread > 37) return;  // premature end: keep the zeros (mcu.rs:206-218)
            for (int col = 0; col < mw; col++) {
                for (Component &c : comp) {
                    const size_t m = (size_t)r * mw + col;
                    size_t idx[4];
                    int nb = 0;
                    if (c.h == 1 && c.v == 1) {
                        idx[nb++] = m;
                    } else if (c.h == 2 && c.v == 2) {
                        const size_t m2 = (size_t)r * mw * 2 + col;
                        idx[nb++] = 2 * m2;
                        idx[nb++] = 2 * m2 + 1;
                        idx[nb++] = 2 * m2 + (size_t)mw * 2;
                        idx[nb++] = 2 * m2 + (size_t)mw * 2 + 1;
                    } else {  // (2, 1) and, as the fork places it, (1, 2)
                        idx[nb++] = 2 * m;
                        idx[nb++] = 2 * m + 1;
                    }
                    for (int b = 0; b < nb; b++) {
                        if ((idx[b] + 1) * 64 > c.coef.size()) continue;  // empty slot: no bits read
                        decode_block_sequential(br, c, &c.coef[idx[b] * 64]);
                    }
                }
                after_sequential_mcu(br, todo);
            }
        }
    }

    // G.1.2.1 / G.1.2.2 / G.1.2.3 for one block of the current progressive scan.
    void prog_dc_first(BitReader &br, Component &c, int16_t *blk)
    {
        const Huffman &hd = dc[c.td];
        if (!hd.set) throw Error{"missing Huffman table"};
        const int t = decode_symbol(br, hd);
        if (t > 16) throw Error{"bad DC magnitude"};
        const int diff = t ? extend(br.get(t), t) : 0;
        c.dc_pred = (int32_t)((uint32_t)c.dc_pred + (uint32_t)diff);
        blk[0] = (int16_t)((uint16_t)(int16_t)c.dc_pred << al);
    }
    void prog_dc_refine(BitReader &br, int16_t *blk)
    {
        if (br.get(1)) blk[0] = (int16_t)((uint16_t)blk[0] + (uint16_t)(1u << al));
    }
    void prog_ac_first(BitReader &br, Component &c, int16_t *blk)
    {
        if (eobrun > 0) {
            eobrun--;
            return;
        }
        const Huffman &ha = ac[c.ta];
        if (!ha.set) throw Error{"missing Huffman table"};
        for (int k = ss; k <= se;) {
            int fl = 0;
            const int rs = decode_symbol(br, ha, &fl);
            const int r = rs >> 4, s = rs & 15;
            if (s) {
                k += r;
                const int v = extend(br.get(s), s);
                blk[ac_position(k, fl, s, v)] = (int16_t)((uint16_t)(int16_t)v << al);
                k++;
            } else if (r == 15) {
                k += 16;
            } else {
                eobrun = (1 << r) - 1;
                if (r) eobrun += (int)br.get(r);
                break;
            }
        }
    }
    void prog_ac_refine(BitReader &br, Component &c, int16_t *blk)
    {
        const int p1 = 1 << al, m1 = -(1 << al);
        int k = ss;
        auto correct = [&](int16_t &x) {
            if (br.get(1) && (x & p1) == 0) x = (int16_t)(x + (x >= 0 ? p1 : m1));
        };
        if (eobrun == 0) {
            const Huffman &ha = ac[c.ta];
            if (!ha.set) throw Error{"missing Huffman table"};
            // k <= se at the top of every pass (bitstream.rs:536-622).
            while (k <= se) {
                const int rs = decode_symbol(br, ha);
                int r = rs >> 4, s = rs & 15;
                if (s) {
                    if (s != 1) throw Error{"bad refinement magnitude"};
                    s = br.get(1) ? p1 : m1;
                } else if (r != 15) {
                    eobrun = 1 << r;
                    if (r) eobrun += (int)br.get(r);
                    break;
                }
                // Correct the nonzero coefficients passed over, up to the r-th zero one.  A run
                // that would pass the end of the band stops at its last position, and a new
                // coefficient then goes there (bitstream.rs:579-616).
                for (;;) {
                    int16_t &x = blk[kNatural[k]];
                    if (x != 0) {
                        correct(x);
                    } else if (--r < 0) {
                        break;
                    }
                    if (k == se) break;
                    k++;
                }
                if (s) blk[kNatural[k]] = (int16_t)s;
                k++;
            }
        }
        if (eobrun > 0) {
            for (; k <= se; k++) {
                int16_t &x = blk[kNatural[k]];
                if (x != 0) correct(x);
            }
            eobrun--;
        }
    }

    void decode_progressive_scan(BitReader &br)
    {
        int todo = restart_interval;
        auto tick = [&]() { after_progressive_unit(br, todo); };
        if (scomp.size() == 1) {
            // One component: its blocks in raster order (mcu_prog.rs:236-330).
            Component &c = comp[scomp[0]];
            int bw, bh;
            if ((scomp[0] == 0 && (c.h != 1 || c.v != 1)) || !interleaved) {
                bw = (width + 7) / 8;
                bh = (height + 7) / 8;
            } else {
                bw = mcu_w;
                bh = mcu_h;
            }
            const int stride = (int)(c.rounded_w / 8);
            for (int i = 0; i < bh; i++)
                for (int j = 0; j < bw; j++) {
                    const size_t at = (size_t)64 * ((size_t)i * stride + j);
                    if (at + 64 > c.coef.size()) throw Error{"block outside the component plane"};
                    int16_t *blk = &c.coef[at];
                    if (ss == 0) {
                        if (ah == 0) prog_dc_first(br, c, blk);
                        else prog_dc_refine(br, blk);
                    } else {
                        if (ah == 0) prog_ac_first(br, c, blk);
                        else prog_ac_refine(br, c, blk);
                    }
                    tick();
                }
        } else {
            // Interleaved DC scan (mcu_prog.rs:335-410).
            const int mw = interleaved ? mcu_w : (width + 7) / 8;
            const int mh = interleaved ? mcu_h : (height + 7) / 8;
            for (int i = 0; i < mh; i++)
                for (int j = 0; j < mw; j++) {
                    for (int k : scomp) {
                        Component &c = comp[k];
                        const int stride = (int)(c.rounded_w / 8);
                        for (int vs = 0; vs < c.v; vs++)
                            for (int hs = 0; hs < c.h; hs++) {
                                const size_t at = (size_t)64 * ((size_t)(i * c.v + vs) * stride + j * c.h + hs);
                                if (at + 64 > c.coef.size()) throw Error{"block outside the component plane"};
                                if (ah == 0) prog_dc_first(br, c, &c.coef[at]);
                                else prog_dc_refine(br, &c.coef[at]);
                            }
                    }
                    tick();
                }
        }
    }

    void run()
    {
        if (n < 2 || d[0] != 0xFF || d[1] != 0xD8) throw Error{"not a JPEG (no SOI marker)"};
        pos = 2;
        bool done = false, decoded_any = false;
        while (!done) {
            const int m = next_marker();
            if (m < 0) {
                if (decoded_any) break;
                throw Error{"no scan before the end of data"};
            }
            switch (m) {
            case 0xD8: break;
            case 0xD9: done = true; break;
            case 0xDB: parse_dqt(); break;
            case 0xC4: parse_dht(); break;
            case 0xDD:
                if (u16() != 4) throw Error{"bad DRI length"};
                restart_interval = u16();
                break;
            case 0xC0:
            case 0xC1:
            case 0xC2: parse_sof(m); break;
            case 0xC3:
            case 0xC5:
            case 0xC6:
            case 0xC7:
            case 0xC9:
            case 0xCA:
            case 0xCB:
            case 0xCD:
            case 0xCE:
            case 0xCF: throw Error{"unsupported JPEG process (lossless, hierarchical or arithmetic coding)"};
            case 0xDA: {
                if (!seen_sof) throw Error{"scan before the frame header"};
                parse_sos();
                if (!decoded_any) setup();
                BitReader br;
                br.d = d;
                br.n = n;
                br.pos = pos;
                for (Component &c : comp) c.dc_pred = 0;
                eobrun = 0;
                if (progressive) {
                    decode_progressive_scan(br);
                } else {
                    decode_sequential(br);
                    decoded_any = true;
                    done = true;  // sequential: the first scan only (mcu.rs:80-285)
                    break;
                }
                decoded_any = true;
                // resume marker parsing after the scan's data
                pos = br.pos;
                if (br.marker >= 0) {
                    // the reader consumed the marker bytes: step back so next_marker sees them
                    pos = br.pos >= 2 ? br.pos - 2 : 0;
                }
                break;
            }
            default:
                if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE || m == 0xDC || m == 0xDE || m == 0xDF) {
                    const int len = u16();
                    if (len < 2 || pos + (size_t)(len - 2) > n) throw Error{"bad segment length"};
                    pos += (size_t)(len - 2);
                } else if (m >= 0xD0 && m <= 0xD7) {
                    // stray restart marker between scans
                } else {
                    throw Error{"unexpected marker"};
                }
            }
        }
        if (!decoded_any) throw Error{"no scan in the data"};
    }
};

}  // namespace

struct artefact_jpeg {
    uint32_t real_w = 0, real_h = 0;
    std::vector<Component> comp;
    std::vector<std::vector<uint16_t>> quant;
    int hmax = 1, vmax = 1;
};

extern "C" {

int artefact_jpeg_decode(const uint8_t *data, size_t size, artefact_jpeg **out, char *err, size_t err_len)
{
    if (err && err_len) err[0] = 0;
    if (!out) return ARTEFACT_E_ARGUMENT;
    *out = nullptr;
    if (!data) return ARTEFACT_E_ARGUMENT;
    artefact_jpeg *j = new (std::nothrow) artefact_jpeg;
    if (!j) return ARTEFACT_E_HIP;
    try {
        Decoder dec(data, size);
        dec.run();
        j->real_w = (uint32_t)dec.width;
        j->real_h = (uint32_t)dec.height;
        // The plane geometry is fixed before a grey reset, so recompute the factors from the
        // frame's original sampling: max / own over the planes just built.
        uint32_t maxw = 0, maxh = 0;
        for (const Component &c : dec.comp) {
            maxw = std::max(maxw, c.rounded_w);
            maxh = std::max(maxh, c.rounded_h);
        }
        j->comp = std::move(dec.comp);
        for (Component &c : j->comp) {
            c.h = (int)(maxw / c.rounded_w);
            c.v = (int)(maxh / c.rounded_h);
            j->quant.emplace_back(dec.qt[c.tq], dec.qt[c.tq] + 64);
        }
    } catch (const Error &e) {
        if (err && err_len) {
            strncpy(err, e.msg.c_str(), err_len - 1);
            err[err_len - 1] = 0;
        }
        delete j;
        return ARTEFACT_E_JPEG;
    } catch (const std::bad_alloc &) {
        delete j;
        return ARTEFACT_E_HIP;
    }
    *out = j;
    return ARTEFACT_OK;
}

uint32_t artefact_jpeg_nchannel(const artefact_jpeg *j) { return j ? (uint32_t)j->comp.size() : 0; }

int artefact_jpeg_size(const artefact_jpeg *j, uint32_t *real_w, uint32_t *real_h)
{
    if (!j || !real_w || !real_h) return ARTEFACT_E_ARGUMENT;
    *real_w = j->real_w;
    *real_h = j->real_h;
    return ARTEFACT_OK;
}

int artefact_jpeg_channels(const artefact_jpeg *j, artefact_channel *out)
{
    if (!j || !out) return ARTEFACT_E_ARGUMENT;
    for (size_t i = 0; i < j->comp.size(); i++) {
        const Component &c = j->comp[i];
        out[i].rounded_w = c.rounded_w;
        out[i].rounded_h = c.rounded_h;
        out[i].block_w = c.rounded_w / 8;
        out[i].block_h = c.rounded_h / 8;
        out[i].h_samp = (uint32_t)c.h;
        out[i].v_samp = (uint32_t)c.v;
        out[i].coefs = c.coef.data();
        out[i].quant = j->quant[i].data();
    }
    return ARTEFACT_OK;
}

void artefact_jpeg_free(artefact_jpeg *j) { delete j; }

}  // extern "C"
