// Host-side session plan of libhbem: triangle records, coordinate-key ids
// for mirror images, the singular (touching) pair list and the
// Sauter-Schwab rule tables.  Built once per session, uploaded once.
//
// The pair enumeration reproduces the reference's singular pair set:
//   - no symmetry: pairs sharing >= 1 vertex index
//     (buildRealDuffyPairList, main.swift:2256-2299);
//   - symmetry: every (test mask, trial mask) combination whose mirrored
//     triangles share a vertex by 1e-6 coordinate key, (0,0) by index
//     (buildSymmetryDuffyPairList main.swift:2301-2380,
//      imageSharedLocalIds main.swift:2152-2193).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <unordered_set>

#include "hbem_internal.h"

namespace hbem {
namespace {

struct Key {
    int64_t x, y, z;
    bool operator<(const Key &o) const {
        if (x != o.x) return x < o.x;
        if (y != o.y) return y < o.y;
        return z < o.z;
    }
    bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
};

// 1e-6 coordinate quantisation, round half away from zero (coordinateKey,
// main.swift:2107-2112).
Key make_key(double x, double y, double z) {
    return Key{std::llround(x / 1.0e-6), std::llround(y / 1.0e-6), std::llround(z / 1.0e-6)};
}

int image_mask_list(int symmetry, int *masks) {
    switch (symmetry) {
    case 1: masks[0] = 1; return 1;
    case 2: masks[0] = 2; return 1;
    case 4: masks[0] = 4; return 1;
    case 3: masks[0] = 1; masks[1] = 2; masks[2] = 3; return 3;
    default: return 0;
    }
}

bool mask_odd(int m) { return __builtin_popcount(static_cast<unsigned>(m)) & 1; }

// Image-local -> original-local vertex order (main.swift:2118-2143).
int image_local(int local, int mask) {
    if (!mask_odd(mask)) return local;
    return local == 1 ? 2 : (local == 2 ? 1 : local);
}

// One touching (test image, trial image) combination.  tl*/sl* are the
// image-local indices of the shared vertices (kinds 2 and 3).
struct PairRec {
    int test, trial, tm, sm, kind, tl1, tl2, sl1, sl2;
};

struct Shared {
    int count = 0, tl1 = 0, tl2 = 0, sl1 = 0, sl2 = 0;
    void add(int t, int s) {
        ++count;
        if (count == 1) {
            tl1 = t;
            sl1 = s;
        } else {
            tl2 = t;
            sl2 = s;
        }
    }
};

}  // namespace

int duffy_rule(int kind, std::vector<double> *out) {
    // Gauss-Legendre order 4 on [0, 1] (gaussRule1D4, main.swift:2626-2635).
    const double s65 = std::sqrt(6.0 / 5.0);
    const double x1 = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * s65) / 2.0;
    const double x2 = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * s65) / 2.0;
    const double w1 = (18.0 + std::sqrt(30.0)) / 72.0;
    const double w2 = (18.0 - std::sqrt(30.0)) / 72.0;
    const double xs[4] = {0.5 - x2, 0.5 - x1, 0.5 + x1, 0.5 + x2};
    const double ws[4] = {w2, w1, w1, w2};
    int n = 0;
    auto put = [&](double a1, double a2, double b1, double b2, double w) {
        // (x - y, y) form of the region maps (appendDuffy, main.swift:2637-2650)
        out->push_back(a1 - a2);
        out->push_back(a2);
        out->push_back(b1 - b2);
        out->push_back(b2);
        out->push_back(w);
        ++n;
    };
    // Sauter-Schwab region maps (duffyRule, main.swift:2652-2699).
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c)
                for (int d = 0; d < 4; ++d) {
                    const double xi = xs[b], e1 = xs[a], e2 = xs[c], e3 = xs[d];
                    const double e12 = e1 * e2, e123 = e12 * e3;
                    const double base = ws[a] * ws[b] * ws[c] * ws[d];
                    if (kind == 1) {
                        const double w = base * xi * xi * xi * e1 * e1 * e2;
                        put(xi, xi * (1 - e1 + e12), xi * (1 - e123), xi * (1 - e1), w);
                        put(xi * (1 - e123), xi * (1 - e1), xi, xi * (1 - e1 + e12), w);
                        put(xi, xi * (e1 - e12 + e123), xi * (1 - e12), xi * (e1 - e12), w);
                        put(xi * (1 - e12), xi * (e1 - e12), xi, xi * (e1 - e12 + e123), w);
                        put(xi * (1 - e123), xi * (e1 - e123), xi, xi * (e1 - e12), w);
                        put(xi, xi * (e1 - e12), xi * (1 - e123), xi * (e1 - e123), w);
                    } else if (kind == 2) {
                        const double w = base * xi * xi * xi * e1 * e1;
                        put(xi, xi * e1 * e3, xi * (1 - e12), xi * e1 * (1 - e2), w);
                        put(xi, xi * e1, xi * (1 - e123), xi * e1 * e2 * (1 - e3), w * e2);
                        put(xi * (1 - e12), xi * e1 * (1 - e2), xi, xi * e123, w * e2);
                        put(xi * (1 - e123), xi * e12 * (1 - e3), xi, xi * e1, w * e2);
                        put(xi * (1 - e123), xi * e1 * (1 - e2 * e3), xi, xi * e12, w * e2);
                    } else {
                        const double w = base * xi * xi * xi * e2;
                        put(xi, xi * e1, xi * e2, xi * e2 * e3, w);
                        put(xi * e2, xi * e2 * e3, xi, xi * e1, w);
                    }
                }
    return n;
}

namespace {

// Shared-vertex remap of a rule point onto the image-local reference
// triangle (remapSingular, main.swift:2709-2735): kind 2 sends the rule's
// vertices (0, 1, 2) to image-local (l1, l2, the third vertex); kind 3
// moves the shared vertex l1 to the rule origin.
void remap_point(double x, double y, int kind, int l1, int l2, double &ox, double &oy) {
    if (kind == 2) {
        const int vc = 3 - l1 - l2;
        const double ax = (l1 == 1), ay = (l1 == 2);
        const double bx = (l2 == 1), by = (l2 == 2);
        const double cx = (vc == 1), cy = (vc == 2);
        ox = ax + x * (bx - ax) + y * (cx - ax);
        oy = ay + x * (by - ay) + y * (cy - ay);
    } else if (kind == 3 && l1 == 1) {
        ox = 1.0 - x - y;
        oy = y;
    } else if (kind == 3 && l1 == 2) {
        ox = x;
        oy = 1.0 - x - y;
    } else {
        ox = x;
        oy = y;
    }
}

// Affine map of rule coordinates onto one side's mirrored triangle:
// point(x, y) = p0 + x e1 + y e2, and loc[a] = original local vertex at rule
// vertex a.  Odd masks swap the image parameters (image local order
// (0, 2, 1), main.swift:2118-2143).
struct SideMap {
    double p0[3], e1[3], e2[3];
    int loc[3];
};

SideMap side_map(int nv, int T, const float *vtx, const int32_t *tri, int t, int mask, int kind,
                 int l1, int l2) {
    double v[3][3];
    for (int l = 0; l < 3; ++l)
        for (int c = 0; c < 3; ++c) {
            const double x = vtx[c * nv + tri[l * T + t]];
            v[l][c] = (mask & (1 << c)) ? -x : x;
        }
    const bool odd = __builtin_popcount(static_cast<unsigned>(mask)) & 1;
    SideMap m;
    double at[3][3];
    const double rx[3] = {0.0, 1.0, 0.0}, ry[3] = {0.0, 0.0, 1.0};
    for (int a = 0; a < 3; ++a) {
        double u, w;
        remap_point(rx[a], ry[a], kind, l1, l2, u, w);
        if (odd) std::swap(u, w);
        m.loc[a] = (u > 0.5) ? 1 : (w > 0.5 ? 2 : 0);
        for (int c = 0; c < 3; ++c) at[a][c] = v[m.loc[a]][c];
    }
    for (int c = 0; c < 3; ++c) {
        m.p0[c] = at[0][c];
        m.e1[c] = at[1][c] - at[0][c];
        m.e2[c] = at[2][c] - at[0][c];
    }
    return m;
}

// Groups the full singular pair list into classes (see Plan::sg_f) and
// writes the per-class records and the per-point rule tables.
void build_singular_classes(int nv, int T, const float *vtx, const int32_t *tri,
                            const int32_t *l2g, const float *areas, const float *normals,
                            const std::vector<PairRec> &pairs, Plan *P) {
    struct Cls {
        int rep;     // index into pairs
        int weight;  // members
    };
    // Class key: (test, trial, tm^sm, kind) and, for coincident and edge
    // pairs, the mask parity and the shared locals.  The vertex rule gives
    // the same block under the odd-mask swap of both triangles (its point
    // set maps onto itself); the coincident and edge rules do not, so their
    // classes keep one representative per parity.  Triangle indices fit 24
    // bits (build_plan rejects larger meshes).
    std::unordered_map<uint64_t, int> index;
    index.reserve(pairs.size());
    std::vector<Cls> cls[3];
    for (size_t i = 0; i < pairs.size(); ++i) {
        const PairRec &p = pairs[i];
        uint64_t key = (static_cast<uint64_t>(p.test) << 38) | (static_cast<uint64_t>(p.trial) << 14) |
                       (static_cast<uint64_t>(p.tm ^ p.sm) << 11) | (static_cast<uint64_t>(p.kind) << 9);
        if (p.kind != 3) {
            const uint64_t par = __builtin_popcount(static_cast<unsigned>(p.tm)) & 1;
            key |= (par << 8) | (static_cast<uint64_t>(p.tl1) << 6) | (static_cast<uint64_t>(p.tl2) << 4) |
                   (static_cast<uint64_t>(p.sl1) << 2) | static_cast<uint64_t>(p.sl2);
        }
        const uint64_t full = key;
        auto it = index.find(full);
        if (it == index.end()) {
            index.emplace(full, static_cast<int>(cls[p.kind - 1].size()));
            cls[p.kind - 1].push_back(Cls{static_cast<int>(i), 1});
        } else {
            cls[p.kind - 1][it->second].weight++;
        }
    }
    int np = 0;
    int64_t nclasses = 0;
    for (int k = 0; k < 3; ++k) {
        P->sg_waves[k] = static_cast<int>((cls[k].size() + 63) / 64);
        np += 64 * P->sg_waves[k];
        nclasses += static_cast<int64_t>(cls[k].size());
    }
    P->sg_np = np;
    P->sg_classes = nclasses;
    P->sg_f.assign(static_cast<size_t>(kSgF) * np, 0.0f);
    P->sg_i.assign(static_cast<size_t>(kSgI) * np, 0);
    P->sg_adj.assign(static_cast<size_t>(kSgAdj) * np, 0);
    for (int p = 0; p < np; ++p) P->sg_i[static_cast<size_t>(6) * np + p] = -1;
    int base = 0;
    for (int k = 0; k < 3; ++k) {
        for (size_t c = 0; c < cls[k].size(); ++c) {
            const PairRec &pr = pairs[cls[k][c].rep];
            const int slot = base + static_cast<int>(c);
            const SideMap tmap = side_map(nv, T, vtx, tri, pr.test, pr.tm, pr.kind, pr.tl1, pr.tl2);
            const SideMap smap = side_map(nv, T, vtx, tri, pr.trial, pr.sm, pr.kind, pr.sl1, pr.sl2);
            float *f = P->sg_f.data();
            auto put = [&](int field, double v) { f[static_cast<size_t>(field) * np + slot] = static_cast<float>(v); };
            for (int d = 0; d < 3; ++d) {
                put(0 + d, smap.p0[d] - tmap.p0[d]);
                put(3 + d, tmap.e1[d]);
                put(6 + d, tmap.e2[d]);
                put(9 + d, smap.e1[d] - tmap.e1[d]);
                put(12 + d, smap.e2[d] - tmap.e2[d]);
                const double nrm = normals[d * T + pr.trial];
                put(15 + d, (pr.sm & (1 << d)) ? -nrm : nrm);
            }
            const double w = cls[k][c].weight;
            put(18, 4.0 * static_cast<double>(areas[pr.test]) * static_cast<double>(areas[pr.trial]) * w);
            const bool mass = pr.kind == 1 && pr.test == pr.trial && pr.tm == pr.sm;
            put(19, mass ? -0.5 * static_cast<double>(areas[pr.test]) * w : 0.0);
            int32_t *ii = P->sg_i.data();
            for (int a = 0; a < 3; ++a) {
                ii[static_cast<size_t>(a) * np + slot] = l2g[pr.test * 3 + tmap.loc[a]];
                ii[static_cast<size_t>(3 + a) * np + slot] = l2g[pr.trial * 3 + smap.loc[a]];
            }
            ii[static_cast<size_t>(6) * np + slot] = pr.trial;
            int32_t *aj = P->sg_adj.data();
            for (int a = 0; a < 3; ++a) {
                aj[static_cast<size_t>(a) * np + slot] = tri[tmap.loc[a] * T + pr.test];
                aj[static_cast<size_t>(3 + a) * np + slot] = tri[smap.loc[a] * T + pr.trial];
            }
            aj[static_cast<size_t>(6) * np + slot] = pr.tm | (pr.sm << 4);
            aj[static_cast<size_t>(7) * np + slot] = pr.test;
        }
        base += 64 * P->sg_waves[k];
    }

    P->sg_rule.clear();
    int off = 0;
    std::vector<double> pts;
    for (int kind = 1; kind <= 3; ++kind) {
        pts.clear();
        const int cnt = duffy_rule(kind, &pts);
        P->sg_rule_offset[kind - 1] = off;
        P->sg_rule_count[kind - 1] = cnt;
        off += cnt;
        for (int i = 0; i < cnt; ++i) {
            const double x = pts[5 * i], y = pts[5 * i + 1], xs = pts[5 * i + 2], ys = pts[5 * i + 3];
            const double w = pts[5 * i + 4];
            const double rho[3] = {1.0 - x - y, x, y}, rhos[3] = {1.0 - xs - ys, xs, ys};
            float row[kSgRule];
            row[0] = static_cast<float>(xs - x);
            row[1] = static_cast<float>(ys - y);
            row[2] = static_cast<float>(xs);
            row[3] = static_cast<float>(ys);
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) row[4 + 3 * a + b] = static_cast<float>(w * rho[a] * rhos[b]);
                row[13 + a] = static_cast<float>(w * rho[a]);
            }
            P->sg_rule.insert(P->sg_rule.end(), row, row + kSgRule);
        }
    }

    // 6 x 6 tensor rule per kind, 64 rows (36 points, 28 zero-weight copies
    // of the first point)
    P->sg_rule_reg.clear();
    for (int kind = 0; kind < 3; ++kind) {
        P->sg_rule_reg_offset[kind] = 64 * kind;
        P->sg_rule_reg_count[kind] = 64;
        for (int p = 0; p < 64; ++p) {
            const int a6 = p < 36 ? p / 6 : 0, b6 = p < 36 ? p % 6 : 0;
            const double x = kRule6X[a6], y = kRule6Y[a6], xs = kRule6X[b6], ys = kRule6Y[b6];
            const double w = p < 36 ? kRule6W[a6] * kRule6W[b6] : 0.0;
            const double rho[3] = {1.0 - x - y, x, y}, rhos[3] = {1.0 - xs - ys, xs, ys};
            float row[kSgRule];
            row[0] = static_cast<float>(xs - x);
            row[1] = static_cast<float>(ys - y);
            row[2] = static_cast<float>(xs);
            row[3] = static_cast<float>(ys);
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) row[4 + 3 * a + b] = static_cast<float>(w * rho[a] * rhos[b]);
                row[13 + a] = static_cast<float>(w * rho[a]);
            }
            P->sg_rule_reg.insert(P->sg_rule_reg.end(), row, row + kSgRule);
        }
    }
}

}  // namespace

std::string build_plan(int nv, int T, int n, int symmetry, const float *vtx, const int32_t *tri,
                       const int32_t *l2g, const float *areas, const float *normals,
                       const int32_t *tags, Plan *P) {
    if (nv <= 0 || T <= 0 || n <= 0) return "geometry must have vertices, triangles and dofs";
    if (T >= (1 << 24)) return "at most 16777215 triangles are supported";
    if (!(symmetry == 0 || symmetry == 1 || symmetry == 2 || symmetry == 3 || symmetry == 4))
        return "symmetry code must be 0, 1, 2, 3 or 4";
    P->n_vertices = nv;
    P->n_triangles = T;
    P->n_dofs = n;
    P->symmetry = symmetry;
    int img[3];
    int ni = image_mask_list(symmetry, img);
    P->n_masks = 1 + ni;
    P->masks[0] = 0;
    for (int i = 0; i < ni; ++i) P->masks[1 + i] = img[i];
    P->row_weight = 1.0f;
    if (symmetry & 1) P->row_weight *= 2.0f;
    if (symmetry & 2) P->row_weight *= 2.0f;
    if (symmetry & 4) P->row_weight *= 2.0f;

    for (int t = 0; t < T; ++t) {
        for (int l = 0; l < 3; ++l) {
            const int v = tri[l * T + t];
            const int d = l2g[t * 3 + l];
            if (v < 0 || v >= nv) return "triangles_3xm contains an out-of-range vertex";
            if (d < 0 || d >= n) return "p1_local2global contains an out-of-range dof";
        }
        if (!(areas[t] > 0.0f) || !std::isfinite(areas[t])) return "triangle areas must be positive";
    }
    for (int i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(vtx[i])) return "vertices must be finite";

    // Coordinate-key ids of every vertex under every mask.
    const int M = P->n_masks;
    std::vector<Key> keys(static_cast<size_t>(M) * nv);
    for (int m = 0; m < M; ++m) {
        const int mask = P->masks[m];
        for (int v = 0; v < nv; ++v) {
            double x = vtx[v], y = vtx[nv + v], z = vtx[2 * nv + v];
            if (mask & 1) x = -x;
            if (mask & 2) y = -y;
            if (mask & 4) z = -z;
            keys[static_cast<size_t>(m) * nv + v] = make_key(x, y, z);
        }
    }
    std::vector<Key> uniq(keys);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int32_t> kid(keys.size());
    for (size_t i = 0; i < keys.size(); ++i)
        kid[i] = static_cast<int32_t>(std::lower_bound(uniq.begin(), uniq.end(), keys[i]) -
                                      uniq.begin());
    auto key_id = [&](int m, int v) { return kid[static_cast<size_t>(m) * nv + v]; };

    // Device triangle order: Morton order of the centroids, so that runs of
    // consecutive triangles (tiles) cover compact patches with few distinct
    // dofs.  perm[p] = original index of device triangle p.
    std::vector<int32_t> perm(T), inv(T);
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        std::vector<double> cen(3 * static_cast<size_t>(T));
        for (int t = 0; t < T; ++t)
            for (int c = 0; c < 3; ++c) {
                const double v = (static_cast<double>(vtx[c * nv + tri[t]]) + vtx[c * nv + tri[T + t]] +
                                  vtx[c * nv + tri[2 * T + t]]) / 3.0;
                cen[3 * t + c] = v;
                lo[c] = std::min(lo[c], v);
                hi[c] = std::max(hi[c], v);
            }
        const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30});
        auto spread = [](uint64_t x) {  // 21 bits -> every third bit
            x &= 0x1fffff;
            x = (x | x << 32) & 0x1f00000000ffffULL;
            x = (x | x << 16) & 0x1f0000ff0000ffULL;
            x = (x | x << 8) & 0x100f00f00f00f00fULL;
            x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
            x = (x | x << 2) & 0x1249249249249249ULL;
            return x;
        };
        std::vector<std::pair<uint64_t, int32_t>> code(T);
        for (int t = 0; t < T; ++t) {
            uint64_t c = 0;
            for (int k = 0; k < 3; ++k) {
                const double u = (cen[3 * t + k] - lo[k]) / span;
                const uint64_t q = static_cast<uint64_t>(std::min(u, 1.0) * 2097151.0);
                c |= spread(q) << k;
            }
            code[t] = {c, t};
        }
        std::sort(code.begin(), code.end());
        for (int p = 0; p < T; ++p) {
            perm[p] = code[p].second;
            inv[code[p].second] = p;
        }
    }
    // Tiles: runs of up to kTileTris consecutive device triangles, cut early
    // when the run's distinct dofs would exceed kTileMaxDofs.
    P->tile_start.clear();
    P->tile_count.clear();
    P->tile_ndofs.clear();
    P->tile_dofs.clear();
    std::vector<int32_t> ldof(3 * static_cast<size_t>(T), 0);
    {
        int p = 0;
        std::vector<int32_t> dofs;
        while (p < T) {
            dofs.clear();
            const int start = p;
            while (p < T && p - start < kTileTris) {
                const int t = perm[p];
                int add = 0;
                for (int l = 0; l < 3; ++l)
                    if (std::find(dofs.begin(), dofs.end(), l2g[t * 3 + l]) == dofs.end()) ++add;
                // the same dof may appear twice in one triangle only for
                // degenerate input, which build_plan rejects via areas
                if ((int)dofs.size() + add > kTileMaxDofs) break;
                for (int l = 0; l < 3; ++l) {
                    const int d = l2g[t * 3 + l];
                    auto it = std::find(dofs.begin(), dofs.end(), d);
                    int pos;
                    if (it == dofs.end()) {
                        pos = (int)dofs.size();
                        dofs.push_back(d);
                    } else {
                        pos = (int)(it - dofs.begin());
                    }
                    ldof[3 * static_cast<size_t>(p) + l] = pos;
                }
                ++p;
            }
            P->tile_start.push_back(start);
            P->tile_count.push_back(p - start);
            P->tile_ndofs.push_back((int)dofs.size());
            dofs.resize(kTileMaxDofs, -1);
            P->tile_dofs.insert(P->tile_dofs.end(), dofs.begin(), dofs.end());
        }
        P->max_tile_dofs = *std::max_element(P->tile_ndofs.begin(), P->tile_ndofs.end());
        const int nt = static_cast<int>(P->tile_start.size());
        P->tile_pairs.clear();
        for (int t = 0; t < nt; ++t)
            for (int s = t; s < nt; ++s) {
                P->tile_pairs.push_back(s);
                P->tile_pairs.push_back(t);
            }
    }

    // Triangle records of device triangle p with its local vertices taken in
    // the order o[0], o[1], o[2].
    auto fill_record = [&](int p, const int o[3], float *f, int32_t *ii) {
        const int t = perm[p];
        ii[19] = t;
        for (int l = 0; l < 3; ++l) ii[20 + l] = ldof[3 * static_cast<size_t>(p) + o[l]];
        const int v0 = tri[o[0] * T + t], v1 = tri[o[1] * T + t], v2 = tri[o[2] * T + t];
        for (int a = 0; a < 6; ++a) {
            const double xi = kRule6X[a], eta = kRule6Y[a], b0 = 1.0 - xi - eta;
            for (int c = 0; c < 3; ++c)
                f[3 * a + c] = static_cast<float>(b0 * vtx[c * nv + v0] + xi * vtx[c * nv + v1] +
                                                  eta * vtx[c * nv + v2]);
        }
        for (int c = 0; c < 3; ++c) f[18 + c] = normals[c * T + t];
        f[21] = areas[t];
        f[22] = 2.0f * areas[t];
        for (int c = 0; c < 3; ++c) {
            f[23 + c] = vtx[c * nv + v0];
            f[26 + c] = vtx[c * nv + v1];
            f[29 + c] = vtx[c * nv + v2];
        }
        for (int l = 0; l < 3; ++l) {
            ii[l] = l2g[t * 3 + o[l]];
            ii[3 + l] = tri[o[l] * T + t];
        }
        for (int m = 0; m < M; ++m)
            for (int l = 0; l < 3; ++l) ii[6 + 3 * m + l] = key_id(m, tri[o[l] * T + t]);
        ii[18] = tags ? tags[t] : 0;
    };
    // Records in device order.
    P->tri_f_aos.assign(static_cast<size_t>(T) * kTriF, 0.0f);
    P->tri_i_aos.assign(static_cast<size_t>(T) * kTriI, -1);
    {
        const int natural[3] = {0, 1, 2};
        for (int p = 0; p < T; ++p)
            fill_record(p, natural, &P->tri_f_aos[static_cast<size_t>(p) * kTriF],
                        &P->tri_i_aos[static_cast<size_t>(p) * kTriI]);
    }
    // Test-side records: each wave's run of tile positions ordered as a path
    // (the next test shares the most dofs with the current one, ties to the
    // one with fewer unvisited edge neighbours, then the lower position), and
    // each test's vertices placed so that dofs shared with the previous test
    // keep their local positions.
    P->tri_t_f.assign(static_cast<size_t>(T) * kTriF, 0.0f);
    P->tri_t_i.assign(static_cast<size_t>(T) * kTriI, -1);
    {
        constexpr int kRun = kTileTris / 4;  // tests per wave
        auto slot = [&](int p, int l) { return ldof[3 * static_cast<size_t>(p) + l]; };
        auto n_shared = [&](int p, int q) {
            int c = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) c += slot(p, a) == slot(q, b);
            return c;
        };
        for (size_t tile = 0; tile < P->tile_start.size(); ++tile) {
            const int t0 = P->tile_start[tile], tc = P->tile_count[tile];
            for (int w0 = 0; w0 < tc; w0 += kRun) {
                const int n = std::min(kRun, tc - w0), base = t0 + w0;
                std::vector<char> used(static_cast<size_t>(n), 0);
                auto free_edges = [&](int i) {
                    int c = 0;
                    for (int j = 0; j < n; ++j)
                        if (j != i && !used[j] && n_shared(base + i, base + j) >= 2) ++c;
                    return c;
                };
                std::vector<int> order;
                int cur = 0;
                for (int i = 1; i < n; ++i)
                    if (free_edges(i) < free_edges(cur)) cur = i;
                for (int k = 0; k < n; ++k) {
                    order.push_back(cur);
                    used[cur] = 1;
                    int best = -1, bs = -1, bf = 0;
                    for (int j = 0; j < n; ++j) {
                        if (used[j]) continue;
                        const int sh = n_shared(base + cur, base + j), fe = free_edges(j);
                        if (sh > bs || (sh == bs && fe < bf)) {
                            best = j;
                            bs = sh;
                            bf = fe;
                        }
                    }
                    cur = best;
                }
                int held[3] = {-1, -1, -1};
                for (int k = 0; k < n; ++k) {
                    const int p = base + order[k];
                    int o[3] = {-1, -1, -1};
                    bool placed[3] = {false, false, false};
                    for (int r = 0; r < 3; ++r)
                        for (int l = 0; l < 3; ++l)
                            if (!placed[l] && o[r] < 0 && held[r] >= 0 && slot(p, l) == held[r]) {
                                o[r] = l;
                                placed[l] = true;
                            }
                    for (int l = 0; l < 3; ++l)
                        for (int r = 0; r < 3 && !placed[l]; ++r)
                            if (o[r] < 0) {
                                o[r] = l;
                                placed[l] = true;
                            }
                    const size_t q = static_cast<size_t>(base + k);
                    fill_record(p, o, &P->tri_t_f[q * kTriF], &P->tri_t_i[q * kTriI]);
                    P->tri_t_i[q * kTriI + 23] = p;
                    for (int r = 0; r < 3; ++r) held[r] = slot(p, o[r]);
                }
            }
        }
    }
    P->tri_f_soa.assign(static_cast<size_t>(T) * kTriF, 0.0f);
    P->tri_i_soa.assign(static_cast<size_t>(T) * kTriI, -1);
    for (int t = 0; t < T; ++t) {
        for (int j = 0; j < kTriF; ++j)
            P->tri_f_soa[static_cast<size_t>(j) * T + t] = P->tri_f_aos[static_cast<size_t>(t) * kTriF + j];
        for (int j = 0; j < kTriI; ++j)
            P->tri_i_soa[static_cast<size_t>(j) * T + t] = P->tri_i_aos[static_cast<size_t>(t) * kTriI + j];
    }

    // Singular pair list.
    std::vector<PairRec> pairs;
    pairs.reserve(static_cast<size_t>(T) * 16 * M);
    auto append = [&](int test, int trial, int tm, int sm, Shared s) {
        int kind;
        if (s.count == 3) {
            kind = 1;
            s.tl1 = 0; s.tl2 = 1; s.sl1 = 0; s.sl2 = 1;
        } else if (s.count == 2) {
            kind = 2;
        } else if (s.count == 1) {
            kind = 3;
        } else {
            return;
        }
        pairs.push_back(PairRec{test, trial, tm, sm, kind, s.tl1, s.tl2, s.sl1, s.sl2});
    };
    auto shared_by_index = [&](int test, int trial) {
        Shared s;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (tri[i * T + test] == tri[j * T + trial]) s.add(i, j);
        return s;
    };
    auto shared_by_key = [&](int test, int trial, int tmi, int smi) {
        Shared s;
        const int tm = P->masks[tmi], sm = P->masks[smi];
        for (int tfl = 0; tfl < 3; ++tfl) {
            const int tk = key_id(tmi, tri[image_local(tfl, tm) * T + test]);
            for (int sfl = 0; sfl < 3; ++sfl)
                if (tk == key_id(smi, tri[image_local(sfl, sm) * T + trial])) s.add(tfl, sfl);
        }
        return s;
    };
    // Incidence: id -> triangles (CSR), per mask.
    const int nid = (symmetry == 0) ? nv : static_cast<int>(uniq.size());
    std::vector<std::vector<int32_t>> start(M), list(M);
    for (int m = 0; m < (symmetry == 0 ? 1 : M); ++m) {
        start[m].assign(static_cast<size_t>(nid) + 1, 0);
        list[m].resize(static_cast<size_t>(T) * 3);
        auto idof = [&](int t, int l) {
            const int v = tri[l * T + t];
            return symmetry == 0 ? v : key_id(m, v);
        };
        for (int t = 0; t < T; ++t)
            for (int l = 0; l < 3; ++l) start[m][idof(t, l) + 1]++;
        for (int i = 0; i < nid; ++i) start[m][i + 1] += start[m][i];
        std::vector<int32_t> cur(start[m].begin(), start[m].end() - 1);
        for (int t = 0; t < T; ++t)
            for (int l = 0; l < 3; ++l) list[m][cur[idof(t, l)]++] = t;
    }
    std::vector<int32_t> stamp(T, 0), cand;
    int st = 0;
    for (int trial = 0; trial < T; ++trial) {
        if (symmetry == 0) {
            ++st;
            cand.clear();
            for (int l = 0; l < 3; ++l) {
                const int v = tri[l * T + trial];
                for (int p = start[0][v]; p < start[0][v + 1]; ++p) {
                    const int test = list[0][p];
                    if (stamp[test] != st) {
                        stamp[test] = st;
                        cand.push_back(test);
                    }
                }
            }
            for (int test : cand) append(test, trial, 0, 0, shared_by_index(test, trial));
            continue;
        }
        for (int smi = 0; smi < M; ++smi)
            for (int tmi = 0; tmi < M; ++tmi) {
                ++st;
                cand.clear();
                for (int l = 0; l < 3; ++l) {
                    const int id = key_id(smi, tri[l * T + trial]);
                    for (int p = start[tmi][id]; p < start[tmi][id + 1]; ++p) {
                        const int test = list[tmi][p];
                        if (stamp[test] != st) {
                            stamp[test] = st;
                            cand.push_back(test);
                        }
                    }
                }
                for (int test : cand) {
                    const Shared s = (tmi == 0 && smi == 0) ? shared_by_index(test, trial)
                                                            : shared_by_key(test, trial, tmi, smi);
                    append(test, trial, P->masks[tmi], P->masks[smi], s);
                }
            }
    }
    P->kind_count[0] = P->kind_count[1] = P->kind_count[2] = 0;
    P->image_pairs = 0;
    std::unordered_set<int64_t> slots;
    slots.reserve(pairs.size() * 4);
    for (const PairRec &pr : pairs) {
        P->kind_count[pr.kind - 1]++;
        if (pr.tm != 0 || pr.sm != 0) P->image_pairs++;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                slots.insert(static_cast<int64_t>(l2g[pr.test * 3 + i]) * n + l2g[pr.trial * 3 + j]);
    }
    P->unique_slots = static_cast<int64_t>(slots.size());
    build_singular_classes(nv, T, vtx, tri, l2g, areas, normals, pairs, P);
    P->vtx.assign(vtx, vtx + 3 * (size_t)nv);
    P->tri.assign(tri, tri + 3 * (size_t)T);
    P->normals.assign(normals, normals + 3 * (size_t)T);
    P->areas.assign(areas, areas + T);
    P->tags.assign(T, 0);
    if (tags) P->tags.assign(tags, tags + T);
    P->l2g.assign(l2g, l2g + 3 * static_cast<size_t>(T));
    return std::string();
}

}  // namespace hbem

namespace hbem {

// Near-field pair selection (buildNearPairList, main.swift:2471-2558) and
// the level-L subdivided rule (subdividedPairBlocks, main.swift:2840-2910).
int64_t build_near_classes(Plan *P, double threshold, int level) {
    const int T = P->n_triangles, nv = P->n_vertices, M = P->n_masks;
    const float *vtx = P->vtx.data();
    const int32_t *tri = P->tri.data();
    struct Metric {
        double c[3], lo[3], hi[3], edge;
        int key[3];
    };
    // coordinate keys per vertex and mask for the touching test
    auto key_of = [&](int v, int mask, int c) {
        double x = vtx[c * nv + v];
        if (mask & (1 << c)) x = -x;
        return std::llround(x / 1.0e-6);
    };
    std::vector<Metric> met(static_cast<size_t>(M) * T);
    for (int m = 0; m < M; ++m)
        for (int t = 0; t < T; ++t) {
            Metric &mt = met[static_cast<size_t>(m) * T + t];
            double p[3][3];
            for (int l = 0; l < 3; ++l)
                for (int c = 0; c < 3; ++c) {
                    const double x = vtx[c * nv + tri[l * T + t]];
                    p[l][c] = (P->masks[m] & (1 << c)) ? -x : x;
                }
            for (int c = 0; c < 3; ++c) {
                mt.c[c] = (p[0][c] + p[1][c] + p[2][c]) / 3.0;
                mt.lo[c] = std::min({p[0][c], p[1][c], p[2][c]});
                mt.hi[c] = std::max({p[0][c], p[1][c], p[2][c]});
            }
            double e = 0.0;
            for (int l = 0; l < 3; ++l) {
                const int o = (l + 1) % 3;
                double d2 = 0.0;
                for (int c = 0; c < 3; ++c) d2 += (p[l][c] - p[o][c]) * (p[l][c] - p[o][c]);
                e = std::max(e, d2);
            }
            mt.edge = std::sqrt(e);
        }
    auto gap = [](double amin, double amax, double bmin, double bmax) {
        if (amax < bmin) return bmin - amax;
        if (bmax < amin) return amin - bmax;
        return 0.0;
    };
    auto touching = [&](int test, int trial, int tm, int sm) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int vt = tri[i * T + test], vs = tri[j * T + trial];
                bool same = true;
                for (int c = 0; c < 3 && same; ++c) same = key_of(vt, tm, c) == key_of(vs, sm, c);
                if (same) return true;
            }
        return false;
    };
    std::unordered_map<uint64_t, int> index;
    std::vector<std::array<int, 4>> cls;  // test, trial, smask (tm = 0), weight
    int64_t pairs = 0;
    for (int trial = 0; trial < T; ++trial)
        for (int smi = 0; smi < M; ++smi)
            for (int tmi = 0; tmi < M; ++tmi) {
                const Metric &b = met[static_cast<size_t>(smi) * T + trial];
                for (int test = 0; test < T; ++test) {
                    const Metric &a = met[static_cast<size_t>(tmi) * T + test];
                    const double cut = threshold * std::max(a.edge, b.edge), cut2 = cut * cut;
                    const double dx = gap(a.lo[0], a.hi[0], b.lo[0], b.hi[0]);
                    const double dy = gap(a.lo[1], a.hi[1], b.lo[1], b.hi[1]);
                    const double dz = gap(a.lo[2], a.hi[2], b.lo[2], b.hi[2]);
                    if (dx * dx + dy * dy + dz * dz > cut2) continue;
                    const double cx = a.c[0] - b.c[0], cy = a.c[1] - b.c[1], cz = a.c[2] - b.c[2];
                    if (cx * cx + cy * cy + cz * cz >= cut2) continue;
                    const int tm = P->masks[tmi], sm = P->masks[smi];
                    if (touching(test, trial, tm, sm)) continue;
                    ++pairs;
                    const uint64_t key = (static_cast<uint64_t>(test) << 36) |
                                         (static_cast<uint64_t>(trial) << 8) |
                                         static_cast<uint64_t>(tm ^ sm);
                    auto it = index.find(key);
                    if (it == index.end()) {
                        index.emplace(key, static_cast<int>(cls.size()));
                        cls.push_back({test, trial, tm ^ sm, 1});
                    } else {
                        cls[it->second][3]++;
                    }
                }
            }
    const int nc = static_cast<int>(cls.size());
    const int np = ((nc + 63) / 64) * 64;
    P->nc_np = np;
    P->nc_pairs = pairs;
    P->nc_f.assign(static_cast<size_t>(kNcF) * std::max(np, 1), 0.0f);
    P->nc_i.assign(static_cast<size_t>(kNcI) * std::max(np, 1), 0);
    for (int c = 0; c < np; ++c) P->nc_i[static_cast<size_t>(6) * np + c] = -1;
    for (int c = 0; c < nc; ++c) {
        const int test = cls[c][0], trial = cls[c][1], sm = cls[c][2], w = cls[c][3];
        double v[3][3], u[3][3];
        for (int l = 0; l < 3; ++l)
            for (int d = 0; d < 3; ++d) {
                v[l][d] = vtx[d * nv + tri[l * T + test]];
                const double y = vtx[d * nv + tri[l * T + trial]];
                u[l][d] = (sm & (1 << d)) ? -y : y;
            }
        auto put = [&](int field, double val) { P->nc_f[static_cast<size_t>(field) * np + c] = static_cast<float>(val); };
        for (int d = 0; d < 3; ++d) {
            put(0 + d, u[0][d] - v[0][d]);
            put(3 + d, v[1][d] - v[0][d]);
            put(6 + d, v[2][d] - v[0][d]);
            put(9 + d, u[1][d] - u[0][d]);
            put(12 + d, u[2][d] - u[0][d]);
            const double nrm = P->normals[d * T + trial];
            put(15 + d, (sm & (1 << d)) ? -nrm : nrm);
        }
        put(18, 4.0 * static_cast<double>(P->areas[test]) * static_cast<double>(P->areas[trial]) * w);
        for (int a = 0; a < 3; ++a) {
            P->nc_i[static_cast<size_t>(a) * np + c] = P->l2g[test * 3 + a];
            P->nc_i[static_cast<size_t>(3 + a) * np + c] = P->l2g[trial * 3 + a];
        }
        P->nc_i[static_cast<size_t>(6) * np + c] = trial;
    }
    // rules: level-L subdivision (midpoint split into 4, recursively), then
    // the plain rule; weights include the sub-triangle area ratio
    std::vector<std::array<double, 6>> sub{{0, 0, 1, 0, 0, 1}};
    for (int l = 0; l < level; ++l) {
        std::vector<std::array<double, 6>> next;
        for (const auto &t : sub) {
            const double ax = t[0], ay = t[1], bx = t[2], by = t[3], cx = t[4], cy = t[5];
            const double abx = (ax + bx) * 0.5, aby = (ay + by) * 0.5;
            const double bcx = (bx + cx) * 0.5, bcy = (by + cy) * 0.5;
            const double cax = (cx + ax) * 0.5, cay = (cy + ay) * 0.5;
            next.push_back({ax, ay, abx, aby, cax, cay});
            next.push_back({abx, aby, bx, by, bcx, bcy});
            next.push_back({cax, cay, bcx, bcy, cx, cy});
            next.push_back({abx, aby, bcx, bcy, cax, cay});
        }
        sub.swap(next);
    }
    P->nc_rule.clear();
    for (const auto &t : sub) {
        const double det = std::fabs((t[2] - t[0]) * (t[5] - t[1]) - (t[3] - t[1]) * (t[4] - t[0]));
        for (int a = 0; a < 6; ++a) {
            P->nc_rule.push_back(static_cast<float>(t[0] + kRule6X[a] * (t[2] - t[0]) + kRule6Y[a] * (t[4] - t[0])));
            P->nc_rule.push_back(static_cast<float>(t[1] + kRule6X[a] * (t[3] - t[1]) + kRule6Y[a] * (t[5] - t[1])));
            P->nc_rule.push_back(static_cast<float>(kRule6W[a] * det));
        }
    }
    P->nc_sub_points = static_cast<int>(sub.size()) * 6;
    for (int a = 0; a < 6; ++a) {
        P->nc_rule.push_back(static_cast<float>(kRule6X[a]));
        P->nc_rule.push_back(static_cast<float>(kRule6Y[a]));
        P->nc_rule.push_back(static_cast<float>(kRule6W[a]));
    }
    P->nc_level = level;
    P->nc_threshold = threshold;
    return nc;
}

}  // namespace hbem
