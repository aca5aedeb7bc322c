// adg_regrid.cpp -- new fine-level box layouts from refinement tags (see
// adg_regrid.hpp).  The steps per coarse level, finest first, follow
// AmrCore::regrid(0, t) -> MakeNewGrids as the reference's Mesh uses it:
// buffer the tags by n_error_buf cells, coarsen them by the blocking factor
// over the refinement ratio, tag the footprint of the new level two above
// (grown by n_proper) so the new levels nest, cluster, refine back, split
// boxes longer than max_grid_size, refine to the fine level.  With level 0
// covering the domain the proper-nesting domain of every level is the
// whole domain, so nesting comes from the footprint tags alone.
#include "adg_regrid.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>

namespace adg {
namespace {

int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

long volume(int D, const IBox& b) {
    long v = 1;
    for (int d = 0; d < D; ++d) v *= (long)(b.hi[d] - b.lo[d] + 1);
    return v;
}

IBox bound(int D, const TagList& p, size_t b, size_t e) {
    IBox r;
    for (int d = 0; d < D; ++d) {
        r.lo[d] = p[b][d];
        r.hi[d] = p[b][d];
    }
    for (size_t i = b + 1; i < e; ++i)
        for (int d = 0; d < D; ++d) {
            r.lo[d] = std::min(r.lo[d], p[i][d]);
            r.hi[d] = std::max(r.hi[d], p[i][d]);
        }
    return r;
}

// Cut of one direction from its signature (tag count per slice).
// status: 0 hole, 1 inflection, 2 bisection, 3 no cut (one slice).
int find_cut(const std::vector<long>& sig, int lo, int* status) {
    const int len = (int)sig.size();
    *status = 3;
    if (len <= 1) return lo;
    const int mid = len / 2;
    int cut = -1;
    for (int i = 0; i < len; ++i)
        if (sig[i] == 0 && (cut < 0 || std::abs(i - mid) < std::abs(cut - mid))) cut = i;
    if (cut >= 0) {
        *status = 0;
        return lo + cut;
    }
    // the strongest sign change of the second difference, 2 slices away
    // from either end, nearest the middle among equals
    std::vector<long> d2((size_t)len, 0);
    for (int i = 1; i < len - 1; ++i) d2[i] = sig[i + 1] - 2 * sig[i] + sig[i - 1];
    long best = -1;
    for (int i = 2; i < len - 2; ++i) {
        const long a = d2[i - 1], b = d2[i];
        if ((a < 0 && b > 0) || (a > 0 && b < 0)) {
            const long jump = std::labs(a - b);
            if (jump > best || (jump == best && std::abs(i - mid) < std::abs(cut - mid))) {
                best = jump;
                cut = i;
            }
        }
    }
    if (best > 2) {
        *status = 1;
        return lo + cut;
    }
    *status = 2;
    return lo + mid;
}

// Merge boxes that share a whole face, until none does: per direction d,
// boxes with equal extents in the other directions are sorted along d and
// runs of abutting boxes joined; the directions are swept until a sweep
// joins nothing.
void simplify(int D, std::vector<IBox>& bl) {
    bool any = true;
    while (any) {
        any = false;
        for (int d = 0; d < D; ++d) {
            auto key_less = [&](const IBox& a, const IBox& b) {
                for (int e = 0; e < D; ++e) {
                    if (e == d) continue;
                    if (a.lo[e] != b.lo[e]) return a.lo[e] < b.lo[e];
                    if (a.hi[e] != b.hi[e]) return a.hi[e] < b.hi[e];
                }
                return a.lo[d] < b.lo[d];
            };
            std::sort(bl.begin(), bl.end(), key_less);
            std::vector<IBox> out;
            out.reserve(bl.size());
            for (const IBox& b : bl) {
                if (!out.empty()) {
                    IBox& o = out.back();
                    bool same = true;
                    for (int e = 0; e < D; ++e)
                        if (e != d && (o.lo[e] != b.lo[e] || o.hi[e] != b.hi[e])) same = false;
                    if (same && o.hi[d] + 1 == b.lo[d]) {
                        o.hi[d] = b.hi[d];
                        any = true;
                        continue;
                    }
                }
                out.push_back(b);
            }
            bl.swap(out);
        }
    }
}

// Split every box longer than chunk in some direction into the fewest
// near-equal pieces no longer than chunk; piece lengths are multiples of
// the largest power of two dividing both the length and chunk.
void max_size(int D, std::vector<IBox>& bl, int chunk) {
    for (int d = 0; d < D; ++d) {
        std::vector<IBox> out;
        for (const IBox& b : bl) {
            const int len = b.hi[d] - b.lo[d] + 1;
            if (len <= chunk) { out.push_back(b); continue; }
            int ratio = 1, bs = chunk, nlen = len;
            while (bs % 2 == 0 && nlen % 2 == 0) { ratio *= 2; bs /= 2; nlen /= 2; }
            const int nblk = nlen / bs + (nlen % bs ? 1 : 0);
            const int sz = nlen / nblk, ex = nlen % nblk;
            int at = b.lo[d];
            for (int k = 0; k < nblk; ++k) {
                IBox p = b;
                p.lo[d] = at;
                at += (k < ex ? sz + 1 : sz) * ratio;
                p.hi[d] = at - 1;
                out.push_back(p);
            }
        }
        bl.swap(out);
    }
}

bool box_less(int D, const IBox& a, const IBox& b) {
    for (int d = D - 1; d >= 0; --d)
        if (a.lo[d] != b.lo[d]) return a.lo[d] < b.lo[d];
    return false;
}

}  // namespace

std::vector<IBox> cluster_tags(int D, const TagList& in, double eff) {
    std::vector<IBox> out;
    if (in.empty()) return out;
    TagList p = in;
    struct Range { size_t b, e; IBox bx; };
    std::vector<Range> work{{0, p.size(), bound(D, p, 0, p.size())}};
    while (!work.empty()) {
        Range c = work.back();
        work.pop_back();
        const long n = (long)(c.e - c.b), vol = volume(D, c.bx);
        if ((double)n >= eff * (double)vol) { out.push_back(c.bx); continue; }
        int status[3] = {3, 3, 3}, cut[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
            std::vector<long> sig((size_t)(c.bx.hi[d] - c.bx.lo[d] + 1), 0);
            for (size_t i = c.b; i < c.e; ++i) sig[(size_t)(p[i][d] - c.bx.lo[d])]++;
            cut[d] = find_cut(sig, c.bx.lo[d], &status[d]);
        }
        int best = 3;
        for (int d = 0; d < D; ++d) best = std::min(best, status[d]);
        if (best == 3) { out.push_back(c.bx); continue; }
        // among the directions with the best kind of cut, the most balanced one
        int dir = -1;
        long bal = -1;
        for (int d = 0; d < D; ++d) {
            if (status[d] != best) continue;
            const long m = std::min(cut[d] - c.bx.lo[d], c.bx.hi[d] - cut[d] + 1);
            if (m > bal) { bal = m; dir = d; }
        }
        auto mid = std::stable_partition(p.begin() + (long)c.b, p.begin() + (long)c.e,
                                         [&](const std::array<int, 3>& q) { return q[dir] < cut[dir]; });
        const size_t m = (size_t)(mid - p.begin());
        if (m == c.b || m == c.e) { out.push_back(c.bx); continue; }
        work.push_back({m, c.e, bound(D, p, m, c.e)});
        work.push_back({c.b, m, bound(D, p, c.b, m)});
    }
    std::sort(out.begin(), out.end(), [&](const IBox& a, const IBox& b) { return box_less(D, a, b); });
    return out;
}

std::vector<std::vector<IBox>> make_new_grids(int D, const int* dom0, const int* per, const AmrSettings& s,
                                              const std::vector<std::vector<IBox>>& old,
                                              const std::vector<TagList>& tags) {
    constexpr int ratio = 2;
    const int bfl = std::max(1, s.blocking_factor / ratio);   // blocking factor in coarse cells
    if (old.empty()) throw std::runtime_error("no level 0");
    const int old_finest = (int)old.size() - 1;
    const int max_crse = std::min(old_finest, s.max_level - 1);
    if ((int)tags.size() < max_crse + 1) throw std::runtime_error("tags missing for a level");
    std::vector<std::vector<IBox>> ng((size_t)std::max(s.max_level, 0) + 1);
    int new_finest = 0;
    for (int levc = max_crse; levc >= 0; --levc) {
        const int levf = levc + 1;
        int dc[3] = {1, 1, 1}, cd[3] = {1, 1, 1};
        long ncell = 1;
        for (int d = 0; d < D; ++d) {
            dc[d] = dom0[d] << levc;
            if (dc[d] % bfl) throw std::runtime_error("domain not divisible by the blocking factor");
            cd[d] = dc[d] / bfl;
            ncell *= cd[d];
        }
        std::vector<uint8_t> tg((size_t)ncell, 0);
        // cell of level levc -> coarsened tag cell; false outside a non-periodic boundary
        auto put = [&](const int* iv) {
            int c[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) {
                int v = iv[d];
                if (v < 0 || v >= dc[d]) {
                    if (!per[d]) return;
                    v = ((v % dc[d]) + dc[d]) % dc[d];
                }
                c[d] = v / bfl;
            }
            long k = 0;
            for (int d = D - 1; d >= 0; --d) k = k * cd[d] + c[d];
            tg[(size_t)k] = 1;
        };
        const int nb = s.n_error_buf;
        for (const auto& t : tags[(size_t)levc]) {
            int o[3] = {0, 0, 0};
            const int n0 = 2 * nb + 1, n1 = D > 1 ? 2 * nb + 1 : 1, n2 = D > 2 ? 2 * nb + 1 : 1;
            for (o[2] = 0; o[2] < n2; ++o[2])
                for (o[1] = 0; o[1] < n1; ++o[1])
                    for (o[0] = 0; o[0] < n0; ++o[0]) {
                        int iv[3] = {0, 0, 0};
                        for (int d = 0; d < D; ++d) iv[d] = t[d] + o[d] - nb;
                        put(iv);
                    }
        }
        // footprint of the new level levf+1 (grown by n_proper at level levf)
        if (levf < new_finest) {
            for (const IBox& b : ng[(size_t)levf + 1]) {
                int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) {
                    lo[d] = floordiv(floordiv(b.lo[d], ratio) - s.n_proper, ratio);
                    hi[d] = floordiv(floordiv(b.hi[d], ratio) + s.n_proper, ratio);
                }
                int iv[3] = {0, 0, 0};
                for (iv[2] = lo[2]; iv[2] <= hi[2]; ++iv[2])
                    for (iv[1] = lo[1]; iv[1] <= hi[1]; ++iv[1])
                        for (iv[0] = lo[0]; iv[0] <= hi[0]; ++iv[0]) put(iv);
            }
        }
        TagList pts;
        for (long k = 0; k < ncell; ++k) {
            if (!tg[(size_t)k]) continue;
            std::array<int, 3> c{0, 0, 0};
            long r = k;
            for (int d = 0; d < D; ++d) { c[d] = (int)(r % cd[d]); r /= cd[d]; }
            pts.push_back(c);
        }
        if (pts.empty()) continue;
        new_finest = std::max(new_finest, levf);
        std::vector<IBox> bl = cluster_tags(D, pts, s.grid_eff);
        for (IBox& b : bl)
            for (int d = 0; d < D; ++d) {
                b.lo[d] *= bfl;
                b.hi[d] = b.hi[d] * bfl + bfl - 1;
            }
        simplify(D, bl);
        max_size(D, bl, std::max(1, s.max_grid_size / ratio));
        for (IBox& b : bl)
            for (int d = 0; d < D; ++d) {
                b.lo[d] *= ratio;
                b.hi[d] = b.hi[d] * ratio + ratio - 1;
            }
        std::sort(bl.begin(), bl.end(), [&](const IBox& a, const IBox& b) { return box_less(D, a, b); });
        ng[(size_t)levf] = bl;
    }
    std::vector<std::vector<IBox>> out;
    for (int l = 1; l <= new_finest; ++l) out.push_back(ng[(size_t)l]);
    return out;
}

}  // namespace adg
