// Batch tables of the tile kernels (k_apply_tiles, tt_apply_mfma.inc): for each tile position in the grid
// (first / interior / last tile in each direction, 64 per tile depth), the tile's elements in batches of 16
// that share no node, with the slot words the kernel reads.  Host code, built once per grid.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "tt_internal.h"

namespace tt {

// Batches of one tile's element box (9 x 9 x EZ elements, box coordinates: element (ex, ey, ez)
// has node-box corner 0 at (ex, ey, ez)) in class order, 16 slots per batch.  Slot word:
//   x = node-box index of corner 0 | owned-corner mask << 22 (0 for padding: a valid address and no
//       owned corner, so its results are discarded);
//   y = owned-node offset of corner 0 (x3; negative or out of range when corner 0 is not owned);
//   z = grid offset of the element from the box's element (0, 0, 0) (padding: the class's first element).
// A corner (cx, cy, cz) is owned when the node (ex+cx, ey+cy, ez+cz) lies in [1,ox]x[1,oy]x[1,oz]: a full
// tile owns 8 x 8 x (EZ - 1) nodes; the last tile in a direction may own fewer (a grid of 129 nodes has 16
// full tiles and one of 1 node in x).  Elements that touch no owned node, and elements outside the grid
// (box element 0 of the first tile in a direction, element ox of the last), are left out of the table,
// so an edge tile contracts only the elements next to its nodes (in 128x64x64 a fifth of the tiles are
// edge tiles; with the full table each cost as much as a full tile).
//
// A batch only needs 16 elements no two of which share a node.  Each class fills whole batches in its
// sorted order; the r = n mod 16 elements left over in each class are packed together, first fit, into
// batches of elements from several classes.  Leftovers taken from the end of each class lie in the same
// top layers and touch each other, so the leftover sets are drawn by a fixed-seed pseudo-random choice,
// the first seed (up to 128) that reaches ceil(n / 16) batches wins, and the table stays the same from run
// to run.  An interior 8x8x4 tile then takes 26 batches instead of 28 (405 elements), an 8x8x2 tile 16
// instead of 19 (243).
struct BatchRng {  // 32-bit linear congruential generator: the same sequence on every host
    uint32_t x;
    explicit BatchRng(uint32_t seed) : x(seed * 2654435761u + 1u) {}
    uint32_t below(uint32_t n) {
        x = 1103515245u * x + 12345u;
        return (x >> 8) % n;
    }
};
template <typename V>
static void batch_shuffle(std::vector<V>& v, BatchRng& r) {
    for (size_t i = v.size(); i > 1; i--) std::swap(v[i - 1], v[r.below((uint32_t)i)]);
}
template <int EZ>
struct ClassBatches {
    static constexpr int EX = 9, EY = 9, HX = 10, HY = 10, X = 8, Y = 8, Z = EZ - 1;
    static constexpr int cnt(int E, int p) { return (E - p + 1) / 2; }
    // the table of a tile owning ox x oy x oz nodes (first[d] / last[d]: the tile is the first / last one
    // in direction d) in a grid of gex x gey elements per layer, appended to `out`; returns its batch count
    static int build(const int (&o)[3], const bool (&first)[3], const bool (&last)[3], int gex, int gey,
                     std::vector<int4>& out) {
        const int E[3] = {EX, EY, EZ};
        int lo[3], hi[3];  // box elements that exist
        for (int d = 0; d < 3; d++) {
            lo[d] = first[d] ? 1 : 0;
            hi[d] = last[d] ? o[d] - 1 : E[d] - 1;
        }
        std::vector<std::vector<int4>> classes(8);
        size_t total = 0;
        for (int c = 0; c < 8; c++) {
            const int px = c & 1, py = (c >> 1) & 1, pz = c >> 2;
            const int nx = cnt(EX, px), ny = cnt(EY, py), nz = cnt(EZ, pz), n = nx * ny * nz;
            std::vector<int4>& cls = classes[c];
            for (int k = 0; k < n; k++) {
                const int ex = px + 2 * (k % nx), ey = py + 2 * ((k / nx) % ny), ez = pz + 2 * (k / (nx * ny));
                if (ex < lo[0] || ex > hi[0] || ey < lo[1] || ey > hi[1] || ez < lo[2] || ez > hi[2]) continue;
                int own = 0;
                for (int a = 0; a < 8; a++) {
                    const int hx = ex + cx_of(a), hy = ey + cy_of(a), hz = ez + cz_of(a);
                    if (hx >= 1 && hx <= o[0] && hy >= 1 && hy <= o[1] && hz >= 1 && hz <= o[2]) own |= 1 << a;
                }
                if (!own) continue;
                const int wx = ((ez * HY + ey) * HX + ex) | own << 22;
                const int wy = ((ex - 1) + X * ((ey - 1) + Y * (ez - 1))) * 3;
                cls.push_back(make_int4(wx, wy, ex + gex * (ey + gey * ez), 0));
            }
            total += cls.size();
        }
        // two elements share a node when their box positions differ by at most 1 in every direction
        auto touch = [](const int4& a, const int4& b) {
            const int ia = a.x & 4095, ib = b.x & 4095;
            return std::abs(ia % HX - ib % HX) <= 1 && std::abs(ia / HX % HY - ib / HX % HY) <= 1 &&
                   std::abs(ia / (HX * HY) - ib / (HX * HY)) <= 1;
        };
        const int lower = (int)((total + 15) / 16);
        std::vector<int4> best;
        int best_n = 1 << 30;
        for (uint32_t seed = 0; seed <= 128 && best_n > lower; seed++) {
            std::vector<int4> full, rest;
            for (int c = 0; c < 8; c++) {
                const std::vector<int4>& cls = classes[c];
                const size_t r = cls.size() % 16;
                std::vector<char> left(cls.size(), 0);
                if (seed == 0) {
                    for (size_t i = cls.size() - r; i < cls.size(); i++) left[i] = 1;
                } else if (r) {
                    std::vector<uint32_t> idx(cls.size());
                    for (size_t i = 0; i < idx.size(); i++) idx[i] = (uint32_t)i;
                    BatchRng rng(seed * 8 + (uint32_t)c);
                    batch_shuffle(idx, rng);
                    for (size_t i = 0; i < r; i++) left[idx[i]] = 1;
                }
                for (size_t i = 0; i < cls.size(); i++) (left[i] ? rest : full).push_back(cls[i]);
            }
            if (seed) {
                BatchRng rng(seed);
                batch_shuffle(rest, rng);
            }
            std::vector<std::vector<int4>> mixed;
            for (const int4& e : rest) {
                bool placed = false;
                for (std::vector<int4>& b : mixed) {
                    if (b.size() == 16) continue;
                    bool ok = true;
                    for (const int4& f : b) ok = ok && !touch(e, f);
                    if (ok) {
                        b.push_back(e);
                        placed = true;
                        break;
                    }
                }
                if (!placed) mixed.push_back({e});
            }
            const int n = (int)(full.size() / 16 + mixed.size());
            if (n < best_n) {
                best_n = n;
                best = full;
                for (std::vector<int4>& b : mixed) {
                    const int pad = b[0].z;  // padding: no owned corner, a valid element's coefficients
                    while (b.size() < 16) b.push_back(make_int4(0, 0, pad, 0));
                    best.insert(best.end(), b.begin(), b.end());
                }
            }
        }
        out.insert(out.end(), best.begin(), best.end());
        return best_n;
    }
};

template <int TZ_>
static void build_tab(const Grid& g, std::vector<int4>& codes, int2* at) {
    // owned extents of the last tile in each direction (a full extent when the node count divides)
    const int full[3] = {8, 8, TZ_};
    const int last_ext[3] = {g.nx - 8 * ((g.nx + 7) / 8 - 1), g.ny - 8 * ((g.ny + 7) / 8 - 1),
                             g.nz - TZ_ * ((g.nz + TZ_ - 1) / TZ_ - 1)};
    for (int c = 0; c < 64; c++) {
        int o[3];
        bool first[3], last[3];
        for (int d = 0; d < 3; d++) {
            first[d] = (c >> (2 * d)) & 1;
            last[d] = (c >> (2 * d + 1)) & 1;
            o[d] = last[d] ? last_ext[d] : full[d];
        }
        at[c].x = (int)codes.size();
        at[c].y = ClassBatches<TZ_ + 1>::build(o, first, last, g.ex, g.ey, codes);
    }
}

void build_batch_tables(const Grid& g, BatchTabStore& st) {
    std::vector<int4> codes;
    std::vector<int2> at(128);
    build_tab<4>(g, codes, at.data());
    build_tab<2>(g, codes, at.data() + 64);
    st.buf.alloc(codes.size());
    st.at.alloc(at.size());
    TT_HIP(hipMemcpy(st.buf.p, codes.data(), codes.size() * sizeof(int4), hipMemcpyHostToDevice));
    TT_HIP(hipMemcpy(st.at.p, at.data(), at.size() * sizeof(int2), hipMemcpyHostToDevice));
    st.t4 = BatchTab{st.buf.p, st.at.p};
    st.t2 = BatchTab{st.buf.p, st.at.p + 64};
}

}  // namespace tt
