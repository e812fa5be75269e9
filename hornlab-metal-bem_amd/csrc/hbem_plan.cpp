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

int duffy_rule_f32(int kind, std::vector<float> *out) {
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
        out->push_back(static_cast<float>(a1 - a2));
        out->push_back(static_cast<float>(a2));
        out->push_back(static_cast<float>(b1 - b2));
        out->push_back(static_cast<float>(b2));
        out->push_back(static_cast<float>(w));
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

std::string build_plan(int nv, int T, int n, int symmetry, const float *vtx, const int32_t *tri,
                       const int32_t *l2g, const float *areas, const float *normals,
                       const int32_t *tags, Plan *P) {
    if (nv <= 0 || T <= 0 || n <= 0) return "geometry must have vertices, triangles and dofs";
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
    }

    // Triangle records, in device order.
    P->tri_f_aos.assign(static_cast<size_t>(T) * kTriF, 0.0f);
    P->tri_i_aos.assign(static_cast<size_t>(T) * kTriI, -1);
    for (int p = 0; p < T; ++p) {
        const int t = perm[p];
        float *f = &P->tri_f_aos[static_cast<size_t>(p) * kTriF];
        int32_t *ii = &P->tri_i_aos[static_cast<size_t>(p) * kTriI];
        ii[19] = t;
        for (int l = 0; l < 3; ++l) ii[20 + l] = ldof[3 * static_cast<size_t>(p) + l];
        const int v0 = tri[t], v1 = tri[T + t], v2 = tri[2 * T + t];
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
            ii[l] = l2g[t * 3 + l];
            ii[3 + l] = tri[l * T + t];
        }
        for (int m = 0; m < M; ++m)
            for (int l = 0; l < 3; ++l) ii[6 + 3 * m + l] = key_id(m, tri[l * T + t]);
        ii[18] = tags ? tags[t] : 0;
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
    struct PairRec {
        int test, trial, code, kind;
    };
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
        const int code = tm | (sm << 3) | (kind << 6) | (s.tl1 << 8) | (s.tl2 << 10) |
                         (s.sl1 << 12) | (s.sl2 << 14);
        pairs.push_back(PairRec{test, trial, code, kind});
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
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const PairRec &a, const PairRec &b) { return a.kind < b.kind; });
    P->pair_tt.resize(pairs.size() * 2);
    P->pair_code.resize(pairs.size());
    P->kind_count[0] = P->kind_count[1] = P->kind_count[2] = 0;
    P->image_pairs = 0;
    std::unordered_set<int64_t> slots;
    slots.reserve(pairs.size() * 4);
    for (size_t p = 0; p < pairs.size(); ++p) {
        P->pair_tt[2 * p] = inv[pairs[p].test];
        P->pair_tt[2 * p + 1] = inv[pairs[p].trial];
        P->pair_code[p] = pairs[p].code;
        P->kind_count[pairs[p].kind - 1]++;
        if ((pairs[p].code & 63) != 0) P->image_pairs++;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                slots.insert(static_cast<int64_t>(l2g[pairs[p].test * 3 + i]) * n +
                             l2g[pairs[p].trial * 3 + j]);
    }
    P->unique_slots = static_cast<int64_t>(slots.size());

    P->rule.clear();
    int off = 0;
    for (int kind = 1; kind <= 3; ++kind) {
        const int c = duffy_rule_f32(kind, &P->rule);
        P->rule_offset[kind - 1] = off;
        P->rule_count[kind - 1] = c;
        off += c;
    }
    P->areas.assign(areas, areas + T);
    P->tags.assign(T, 0);
    if (tags) P->tags.assign(tags, tags + T);
    P->l2g.assign(l2g, l2g + 3 * static_cast<size_t>(T));
    return std::string();
}

}  // namespace hbem
