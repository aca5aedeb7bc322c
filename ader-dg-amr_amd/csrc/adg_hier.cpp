// adg_hier.cpp -- builds the flattened hierarchy lists (see adg_hier.hpp).
#include "adg_hier.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace adg {

long Hierarchy::slot_of(int p, const int* iv) const {
    const PatchLayout& P = patch[p];
    long s = 0;
    for (int d = D - 1; d >= 0; --d) s = s * P.g[d] + (iv[d] - P.lo[d] + 1);
    return P.off + s;
}

void Hierarchy::take_storage(Hierarchy& o) {
    auto take = [](auto& dst, auto& src) {
        dst.swap(src);
        dst.clear();
        src.clear();
        src.shrink_to_fit();
    };
    // lists the build sizes and then writes element by element (or refills
    // in parallel) keep their contents, so that sizing them initialises
    // nothing it is about to overwrite
    auto keep = [](auto& dst, auto& src) {
        dst.swap(src);
        src.clear();
        src.shrink_to_fit();
    };
    take(lev, o.lev);
    take(patch, o.patch);
    keep(slot_lev, o.slot_lev);
    keep(valid_list, o.valid_list);
    for (int d = 0; d < 3; ++d) {
        keep(cell_face[d][0], o.cell_face[d][0]);
        keep(cell_face[d][1], o.cell_face[d][1]);
        keep(face_lr[d], o.face_lr[d]);
        keep(face_lev[d], o.face_lev[d]);
        keep(face_tr[d], o.face_tr[d]);
        keep(bcoarse[d], o.bcoarse[d]);
        keep(bfine[d], o.bfine[d]);
        keep(cidx[d], o.cidx[d]);
    }
    take(g_slot, o.g_slot);
    take(g_src, o.g_src);
    take(g_kind, o.g_kind);
    take(g_child, o.g_child);
    take(ad_coarse, o.ad_coarse);
    take(ad_fine, o.ad_fine);
    take(mom_face, o.mom_face);
    take(mom_d, o.mom_d);
    take(mom_kind, o.mom_kind);
    take(mom_sel, o.mom_sel);
    take(mom_lev, o.mom_lev);
    take(rc_slot, o.rc_slot);
    take(rc_start, o.rc_start);
    take(rc_count, o.rc_count);
    take(rc_lev, o.rc_lev);
    take(re_mom, o.re_mom);
    take(re_coef, o.re_coef);
    take(uncovered, o.uncovered);
    take(pred_slots, o.pred_slots);
    take(pred_off, o.pred_off);
    take(copy_ghost, o.copy_ghost);
    take(copy_src, o.copy_src);
    keep(lim_cells, o.lim_cells);
    keep(lim_nbr, o.lim_nbr);
    keep(unc_centre, o.unc_centre);
    // the owner grids keep their per-level storage (clear keeps the inner
    // vectors' capacity only if they stay; they are refilled by assign)
    grid_slot.swap(o.grid_slot);
    grid_patch.swap(o.grid_patch);
    o.grid_slot.clear();
    o.grid_patch.clear();
    for (int k = 0; k < 2; ++k)
        for (int d = 0; d < 3; ++d) mom_tab[k][d].swap(o.mom_tab[k][d]);   // contents kept: all -1
}

long Hierarchy::face_of(int p, int d, const int* iv) const {
    const PatchLayout& P = patch[p];
    long s = 0;
    for (int a = D - 1; a >= 0; --a) s = s * (P.n[a] + (a == d ? 1 : 0)) + (iv[a] - P.lo[a]);
    return P.foff[d] + s;
}

namespace {

struct Grid {  // owner map of one level over its domain: cell -> (patch, slot)
    int D;
    int dom[3];
    std::vector<int> slot, pidx;
    long lin(const int* iv) const {
        long s = 0;
        for (int d = D - 1; d >= 0; --d) s = s * dom[d] + iv[d];
        return s;
    }
};

int floordiv2(int i) { return (i >= 0) ? i / 2 : -((-i + 1) / 2); }

// v sized n and every element set to val, in chunks on the host threads
template <class T>
void fill_par(std::vector<T>& v, size_t n, T val) {
    v.resize(n);
    const long nn = (long)n;
    const int nch = (int)std::min<long>(64, std::max<long>(1, nn >> 20));
    parallel_for(nch, [&](int k) { std::fill(v.begin() + nn * k / nch, v.begin() + nn * (k + 1) / nch, val); });
}

template <class F>
void for_box(int D, const int* lo, const int* hi, F&& f) {
    int iv[3] = {lo[0], lo[1], D > 2 ? lo[2] : 0};
    int h2 = D > 2 ? hi[2] : 0;
    for (iv[2] = D > 2 ? lo[2] : 0; iv[2] <= h2; ++iv[2])
        for (iv[1] = lo[1]; iv[1] <= hi[1]; ++iv[1])
            for (iv[0] = lo[0]; iv[0] <= hi[0]; ++iv[0]) f(iv);
}

}  // namespace

Hierarchy build_hierarchy(int D, const int* dom0, const double* plo, const double* phi,
                          const int* per, int nlev, const int* npatch, const int* boxes, int independent, int conservative_reflux,
                          Hierarchy* reuse) {
    if (D < 2 || D > 3 || nlev < 1 || nlev > 8) throw std::runtime_error("bad dim or level count");
    if (independent && nlev != 1) throw std::runtime_error("independent patches form a single level");
    Hierarchy h;
    if (reuse) h.take_storage(*reuse);
    for (int d = D; d < 3; ++d) {   // directions past D hold nothing
        h.cell_face[d][0].clear();
        h.cell_face[d][1].clear();
        h.face_lr[d].clear();
        h.face_lev[d].clear();
        h.face_tr[d].clear();
        h.bcoarse[d].clear();
        h.bfine[d].clear();
        h.cidx[d].clear();
    }
    h.D = D;
    h.nlev = nlev;
    h.independent = independent != 0;
    for (int d = 0; d < 3; ++d) {
        h.dom0[d] = d < D ? dom0[d] : 1;
        h.plo[d] = d < D ? plo[d] : 0.0;
        h.phi[d] = d < D ? phi[d] : 1.0;
        h.per[d] = d < D ? (per[d] != 0) : 1;
    }
    h.lev.resize(nlev);
    // ---- layout
    int at = 0;
    long soff = 0, foff[3] = {0, 0, 0};
    for (int l = 0; l < nlev; ++l) {
        LevelLayout& L = h.lev[l];
        for (int d = 0; d < D; ++d) {
            L.dom[d] = h.dom0[d] << l;
            L.dx[d] = (h.phi[d] - h.plo[d]) / L.dom[d];
        }
        L.slot_off = soff;
        for (int d = 0; d < D; ++d) L.face_off[d] = foff[d];
        L.p_begin = (int)h.patch.size();
        for (int b = 0; b < npatch[l]; ++b) {
            PatchLayout P;
            P.lev = l;
            for (int d = 0; d < D; ++d) {
                P.lo[d] = boxes[at + d];
                P.hi[d] = boxes[at + D + d];
                if (P.hi[d] < P.lo[d]) throw std::runtime_error("empty box");
                if (P.lo[d] < 0 || P.hi[d] >= L.dom[d]) throw std::runtime_error("box outside domain");
                if (l > 0 && ((P.lo[d] & 1) || !(P.hi[d] & 1)))
                    throw std::runtime_error("fine box not aligned to refinement ratio 2");
                P.n[d] = P.hi[d] - P.lo[d] + 1;
                P.g[d] = P.n[d] + 2;
            }
            at += 2 * D;
            P.off = soff;
            soff += (long)P.g[0] * P.g[1] * P.g[2];
            for (int d = 0; d < D; ++d) {
                long nf = 1;
                for (int a = 0; a < D; ++a) nf *= P.n[a] + (a == d ? 1 : 0);
                P.foff[d] = foff[d];
                foff[d] += nf;
            }
            h.patch.push_back(P);
        }
        L.p_end = (int)h.patch.size();
        L.nslot = soff - L.slot_off;
        for (int d = 0; d < D; ++d) L.nface[d] = foff[d] - L.face_off[d];
    }
    h.NS = soff;
    for (int d = 0; d < D; ++d) h.NF[d] = foff[d];
    if (h.NS >= (1L << 31)) throw std::runtime_error("hierarchy too large for 32-bit slot ids");

    // ---- owner grids (valid cells), overlap check
    std::vector<Grid> G(nlev);
    h.grid_slot.resize(nlev);
    h.grid_patch.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        Grid& g = G[l];
        g.slot.swap(h.grid_slot[l]);   // recycled storage, refilled below
        g.pidx.swap(h.grid_patch[l]);
        g.D = D;
        long nd = 1;
        for (int d = 0; d < 3; ++d) { g.dom[d] = d < D ? h.lev[l].dom[d] : 1; nd *= g.dom[d]; }
        // boxes of a level must not overlap: sweep along x over boxes sorted by lo
        std::vector<int> ord;
        for (int p = h.lev[l].p_begin; p < h.lev[l].p_end; ++p) ord.push_back(p);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return h.patch[a].lo[0] < h.patch[b].lo[0]; });
        for (size_t i = 0; i < ord.size(); ++i)
            for (size_t j = i + 1; j < ord.size() && h.patch[ord[j]].lo[0] <= h.patch[ord[i]].hi[0]; ++j) {
                const PatchLayout &A = h.patch[ord[i]], &B = h.patch[ord[j]];
                bool meet = true;
                for (int d = 0; d < D; ++d)
                    if (A.hi[d] < B.lo[d] || B.hi[d] < A.lo[d]) meet = false;
                if (meet) throw std::runtime_error("overlapping boxes on a level");
            }
        fill_par(g.slot, (size_t)nd, -1);
        fill_par(g.pidx, (size_t)nd, -1);
        parallel_for(h.lev[l].p_end - h.lev[l].p_begin, [&](int i) {
            const int p = h.lev[l].p_begin + i;
            for_box(D, h.patch[p].lo, h.patch[p].hi, [&](const int* iv) {
                const long c = g.lin(iv);
                g.slot[c] = (int)h.slot_of(p, iv);
                g.pidx[c] = p;
            });
        });
    }
    auto wrap = [&](int l, int* iv) {
        for (int d = 0; d < D; ++d) {
            int n = h.lev[l].dom[d];
            if (iv[d] < 0 || iv[d] >= n) {
                if (!h.per[d]) return false;
                iv[d] = ((iv[d] % n) + n) % n;
            }
        }
        return true;
    };
    auto find = [&](int l, const int* iv0) -> long {
        int iv[3] = {iv0[0], iv0[1], D > 2 ? iv0[2] : 0};
        if (!wrap(l, iv)) return -1;
        return G[l].slot[G[l].lin(iv)];
    };
    auto covered = [&](int l, const int* iv0) {  // cell of level l under level l+1
        if (l + 1 >= nlev) return false;
        int iv[3] = {iv0[0], iv0[1], D > 2 ? iv0[2] : 0};
        int f[3] = {2 * iv[0], 2 * iv[1], 2 * iv[2]};
        return G[l + 1].slot[G[l + 1].lin(f)] >= 0;
    };
    // proper nesting: every fine cell's parent must be a valid coarse cell
    for (int l = 1; l < nlev; ++l)
        for (int p = h.lev[l].p_begin; p < h.lev[l].p_end; ++p) {
            int clo[3], chi[3];
            for (int d = 0; d < D; ++d) { clo[d] = h.patch[p].lo[d] >> 1; chi[d] = h.patch[p].hi[d] >> 1; }
            for_box(D, clo, chi, [&](const int* iv) {
                if (find(l - 1, iv) < 0) throw std::runtime_error("fine box not nested in coarse level");
            });
        }

    // ---- per-slot level, valid list, cell faces
    h.slot_lev.resize(h.NS);   // every slot set below
    for (int l = 0; l < nlev; ++l)
        std::fill(h.slot_lev.begin() + h.lev[l].slot_off, h.slot_lev.begin() + h.lev[l].slot_off + h.lev[l].nslot,
                  (uint8_t)l);
    // valid cells in patch order (each patch's range from the cell counts)
    const int np_all = (int)h.patch.size();
    std::vector<long> vbeg((size_t)np_all + 1, 0);
    for (int p = 0; p < np_all; ++p) vbeg[p + 1] = vbeg[p] + (long)h.patch[p].n[0] * h.patch[p].n[1] * h.patch[p].n[2];
    // every valid cell's entries written below, patch by patch
    h.valid_list.resize((size_t)vbeg[np_all]);
    for (int d = 0; d < D; ++d)
        for (int s = 0; s < 2; ++s) h.cell_face[d][s].resize((size_t)vbeg[np_all]);
    for (int l = 0; l < nlev; ++l) {
        h.lev[l].valid_begin = vbeg[h.lev[l].p_begin];
        h.lev[l].valid_end = vbeg[h.lev[l].p_end];
    }
    std::vector<std::vector<int>> unc_p((size_t)np_all);
    parallel_for(np_all, [&](int p) {
        long v = vbeg[p];
        const int l = h.patch[p].lev;
        for_box(D, h.patch[p].lo, h.patch[p].hi, [&](const int* iv) {
            h.valid_list[(size_t)v] = (int)h.slot_of(p, iv);
            for (int d = 0; d < D; ++d) {
                int hv[3] = {iv[0], iv[1], iv[2]};
                hv[d] += 1;
                h.cell_face[d][0][(size_t)v] = (int)h.face_of(p, d, iv);
                h.cell_face[d][1][(size_t)v] = (int)h.face_of(p, d, hv);
            }
            if (!covered(l, iv)) unc_p[(size_t)p].push_back((int)h.slot_of(p, iv));
            ++v;
        });
    });
    for (const std::vector<int>& u : unc_p) h.uncovered.insert(h.uncovered.end(), u.begin(), u.end());
    // ---- faces
    for (int d = 0; d < D; ++d) {
        h.face_lr[d].resize(2 * h.NF[d]);   // every face of every patch written below
        h.face_lev[d].resize(h.NF[d]);
        parallel_for(np_all, [&](int p) {
            const PatchLayout& P = h.patch[p];
            int hi[3] = {P.hi[0], P.hi[1], P.hi[2]};
            hi[d] += 1;
            for_box(D, P.lo, hi, [&](const int* iv) {
                int lv[3] = {iv[0], iv[1], iv[2]};
                lv[d] -= 1;
                long f = h.face_of(p, d, iv);
                h.face_lr[d][2 * f] = (int)h.slot_of(p, lv);
                h.face_lr[d][2 * f + 1] = (int)h.slot_of(p, iv);
                h.face_lev[d][f] = (uint8_t)P.lev;
            });
        });
    }
    // ---- ghost sources (FillPatchSingleLevel / FillPatchTwoLevels semantics),
    // per patch, then concatenated in patch order
    struct GhostList { std::vector<int> slot, src; std::vector<uint8_t> kind, child; };
    std::vector<GhostList> gp((size_t)np_all);
    parallel_for(np_all, [&](int p) {
        GhostList& gl = gp[(size_t)p];
        const PatchLayout& P = h.patch[p];
        const int l = P.lev;
        int glo[3] = {P.lo[0] - 1, P.lo[1] - 1, D > 2 ? P.lo[2] - 1 : 0};
        int ghi[3] = {P.hi[0] + 1, P.hi[1] + 1, D > 2 ? P.hi[2] + 1 : 0};
        for_box(D, glo, ghi, [&](const int* iv) {
            bool inside = true;
            for (int d = 0; d < D; ++d)
                if (iv[d] < P.lo[d] || iv[d] > P.hi[d]) inside = false;
            if (inside) return;
            long s = h.slot_of(p, iv);
            if (h.independent) {   // periodic inside the patch's own domain
                int w[3] = {iv[0], iv[1], D > 2 ? iv[2] : 0};
                for (int d = 0; d < D; ++d) w[d] = P.lo[d] + (((w[d] - P.lo[d]) % P.n[d]) + P.n[d]) % P.n[d];
                gl.slot.push_back((int)s); gl.src.push_back((int)h.slot_of(p, w));
                gl.kind.push_back(0); gl.child.push_back(0);
                return;
            }
            // outside a non-periodic boundary: transmissive (outflow) ghost,
            // the mirror image of the cell across the boundary face, whose
            // modes change sign by (-1)^{k_d} in each mirrored direction d
            int mv[3] = {iv[0], iv[1], D > 2 ? iv[2] : 0}, mask = 0;
            for (int d = 0; d < D; ++d) {
                const int n = h.lev[l].dom[d];
                if (h.per[d]) continue;
                if (iv[d] < 0) { mv[d] = -1 - iv[d]; mask |= 1 << d; }
                if (iv[d] >= n) { mv[d] = 2 * n - 1 - iv[d]; mask |= 1 << d; }
            }
            long src = find(l, mv);
            if (src >= 0) {
                gl.slot.push_back((int)s); gl.src.push_back((int)src);
                gl.kind.push_back((uint8_t)(0 | (mask << 2))); gl.child.push_back(0);
                return;
            }
            if (l == 0) throw std::runtime_error("level-0 ghost cell without a source");
            int civ[3] = {0, 0, 0}, child = 0;
            for (int d = 0; d < D; ++d) {
                civ[d] = floordiv2(mv[d]);
                child += (mv[d] - 2 * civ[d]) << (D - 1 - d);
            }
            long cs = find(l - 1, civ);
            if (cs < 0) throw std::runtime_error("coarse parent of a fine ghost cell is missing");
            gl.slot.push_back((int)s); gl.src.push_back((int)cs);
            gl.kind.push_back((uint8_t)(1 | (mask << 2))); gl.child.push_back((uint8_t)child);
        });
    });
    for (const GhostList& gl : gp) {
        h.g_slot.insert(h.g_slot.end(), gl.slot.begin(), gl.slot.end());
        h.g_src.insert(h.g_src.end(), gl.src.begin(), gl.src.end());
        h.g_kind.insert(h.g_kind.end(), gl.kind.begin(), gl.kind.end());
        h.g_child.insert(h.g_child.end(), gl.child.begin(), gl.child.end());
    }
    // ---- average down groups
    const int nch = 1 << D;
    h.ad_coarse.assign(nlev, {});
    h.ad_fine.assign(nlev, {});
    for (int l = 1; l < nlev; ++l) {
        // a fine patch's groups: its coarse footprint, in patch order
        const int pb = h.lev[l].p_begin, npl = h.lev[l].p_end - pb;
        std::vector<long> gb((size_t)npl + 1, 0);
        for (int i = 0; i < npl; ++i) gb[i + 1] = gb[i] + (vbeg[pb + i + 1] - vbeg[pb + i]) / nch;
        h.ad_coarse[l].assign((size_t)gb[npl], 0);
        h.ad_fine[l].assign((size_t)gb[npl] * nch, 0);
        parallel_for(npl, [&](int i) {
            const int p = pb + i;
            long g = gb[i];
            int clo[3], chi[3];
            for (int d = 0; d < D; ++d) { clo[d] = h.patch[p].lo[d] >> 1; chi[d] = h.patch[p].hi[d] >> 1; }
            for_box(D, clo, chi, [&](const int* iv) {
                h.ad_coarse[l][(size_t)g] = (int)find(l - 1, iv);
                for (int c = 0; c < nch; ++c) {
                    int fv[3] = {2 * iv[0], 2 * iv[1], 2 * iv[2]};
                    for (int d = 0; d < D; ++d) fv[d] += (c >> (D - 1 - d)) & 1;
                    h.ad_fine[l][(size_t)g * nch + c] = (int)h.slot_of(p, fv);
                }
                ++g;
            });
        });
    }
    // ---- C-F face masks (AmrDG_Level.tpp:240-414)
    for (int d = 0; d < D; ++d) {
        fill_par(h.bcoarse[d], (size_t)h.NF[d], (int8_t)0);
        fill_par(h.bfine[d], (size_t)h.NF[d], (int8_t)0);
        fill_par(h.cidx[d], (size_t)h.NF[d], (int8_t)0);
    }
    parallel_for(np_all, [&](int p) {
        const PatchLayout& P = h.patch[p];
        const int l = P.lev;
        // coverage / validity as seen through one ghost layer of this patch
        auto cov_at = [&](const int* iv) -> int {
            int w[3] = {iv[0], iv[1], D > 2 ? iv[2] : 0};
            bool indom = wrap(l, w);
            if (indom) {
                if (G[l].slot[G[l].lin(w)] < 0) return 0;  // not a valid cell of this level
                return covered(l, w) ? 1 : 0;
            }
            int a[3] = {iv[0], iv[1], D > 2 ? iv[2] : 0};  // non-periodic: adjacent valid cell
            for (int dd = 0; dd < D; ++dd) a[dd] = std::min(std::max(a[dd], 0), h.lev[l].dom[dd] - 1);
            return covered(l, a) ? 1 : 0;
        };
        auto valid_at = [&](const int* iv) -> int {
            int w[3] = {iv[0], iv[1], D > 2 ? iv[2] : 0};
            if (!wrap(l, w)) return 1;
            return G[l].slot[G[l].lin(w)] >= 0 ? 1 : 0;
        };
        for (int d = 0; d < D; ++d) {
            int hi[3] = {P.hi[0], P.hi[1], P.hi[2]};
            hi[d] += 1;
            for_box(D, P.lo, hi, [&](const int* iv) {
                int lv[3] = {iv[0], iv[1], iv[2]};
                lv[d] -= 1;
                long f = h.face_of(p, d, iv);
                if (l + 1 < nlev && lv[d] >= P.lo[d]) {  // ownership: skip the patch's lo face
                    int ml = cov_at(lv), mr = cov_at(iv);
                    if (ml != mr) h.bcoarse[d][f] = (ml == 0) ? 1 : -1;
                }
                if (l > 0) {
                    int vl = valid_at(lv), vr = valid_at(iv);
                    if (vl != vr) {
                        h.bfine[d][f] = (vl == 0) ? 1 : -1;
                        int ci = 0, bp = 0;
                        for (int a = 0; a < D; ++a) {
                            if (a == d) continue;
                            if (iv[a] % 2 != 0) ci |= 1 << bp;
                            bp++;
                        }
                        h.cidx[d][f] = (int8_t)ci;
                    }
                }
            });
        }
    });
    // ---- flux-register algebra (FluxRegister CrseAdd(-1) / FineAdd(+1) /
    // Reflux(vol=1) as called in AmrDG.h:1138-1153 and AmrDG_Level.tpp:416-446)
    // (kind, d, face) -> moment index, a face-indexed table per kind and
    // direction (-1: none yet), kept with the hierarchy's storage and reset
    // entry by entry below, so that no build fills a table of every face
    std::vector<int> (&mom_id)[2][3] = h.mom_tab;
    auto moment = [&](int kind, int d, long face, int sel, int lev) {
        std::vector<int>& tab = mom_id[kind][d];
        if (tab.size() < (size_t)h.NF[d]) tab.resize((size_t)h.NF[d], -1);
        if (tab[(size_t)face] >= 0) return tab[(size_t)face];
        int id = (int)h.mom_face.size();
        tab[(size_t)face] = id;
        h.mom_face.push_back((int)face);
        h.mom_d.push_back((uint8_t)d);
        h.mom_kind.push_back((uint8_t)kind);
        h.mom_sel.push_back((uint8_t)sel);
        h.mom_lev.push_back((uint8_t)lev);
        return id;
    };
    // coarse face arrays hit by a coarse face index (with its periodic image along d)
    auto coarse_arrays = [&](int lc, int d, const int* fv0, std::vector<std::pair<int, long>>& out) {
        out.clear();
        int imgs[2] = {fv0[d], -1};
        int nimg = 1;
        if (h.per[d]) {
            if (fv0[d] == 0) imgs[nimg++] = h.lev[lc].dom[d];
            else if (fv0[d] == h.lev[lc].dom[d]) imgs[nimg++] = 0;
        }
        // a face lies in the face array of the patches owning the cell on
        // either side of it (patches in increasing order)
        for (int ii = 0; ii < nimg; ++ii) {
            int fv[3] = {fv0[0], fv0[1], D > 2 ? fv0[2] : 0};
            fv[d] = imgs[ii];
            bool indom = true;
            for (int a = 0; a < D; ++a)
                if (a != d && (fv[a] < 0 || fv[a] >= h.lev[lc].dom[a])) indom = false;
            if (!indom) continue;
            int own[2] = {-1, -1};
            int lv[3] = {fv[0], fv[1], fv[2]};
            lv[d] -= 1;
            if (lv[d] >= 0) own[0] = G[lc].pidx[G[lc].lin(lv)];
            if (fv[d] < h.lev[lc].dom[d]) own[1] = G[lc].pidx[G[lc].lin(fv)];
            if (own[0] > own[1]) std::swap(own[0], own[1]);
            for (int k = 0; k < 2; ++k) {
                if (own[k] < 0 || (k == 1 && own[1] == own[0])) continue;
                out.push_back({own[k], h.face_of(own[k], d, fv)});
            }
        }
    };
    for (int lf = 1; lf < nlev; ++lf) {
        const int lc = lf - 1;
        // Register terms of the coarse faces under each fine patch, found per
        // patch in parallel: for each such face the coarse face arrays it
        // lies in and its terms (coarse side -1, fine faces +1).  Moment ids
        // are then assigned patch by patch in the order the terms were found,
        // so the lists do not depend on the thread count.
        struct Term { long face; int sel; int kind; double coef; };
        struct Rec { int d, a0, a1, t0, t1; };
        struct PatchRegs { std::vector<Rec> rec; std::vector<long> arr; std::vector<Term> term; };
        const int nfp = h.lev[lf].p_end - h.lev[lf].p_begin;
        std::vector<PatchRegs> pr((size_t)nfp);
        parallel_for(nfp, [&](int i) {
            const int B = h.lev[lf].p_begin + i;
            PatchRegs& R = pr[(size_t)i];
            std::vector<std::pair<int, long>> arrs;
            int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { clo[d] = h.patch[B].lo[d] >> 1; chi[d] = h.patch[B].hi[d] >> 1; }
            for (int d = 0; d < D; ++d)
                for (int side = 0; side < 2; ++side) {
                    int tlo[3] = {clo[0], clo[1], clo[2]}, thi[3] = {chi[0], chi[1], chi[2]};
                    tlo[d] = thi[d] = side == 0 ? clo[d] : chi[d] + 1;
                    for_box(D, tlo, thi, [&](const int* fv) {
                        // the coarse cell across this face, outside patch B:
                        // covered by the fine level (another fine patch) or
                        // outside a non-periodic domain, the face has no
                        // register terms (both sides covered, or no fine face
                        // flagged)
                        {
                            int nb[3] = {fv[0], fv[1], D > 2 ? fv[2] : 0};
                            if (side == 0) nb[d] -= 1;
                            if (!wrap(lc, nb) || covered(lc, nb)) return;
                        }
                        const int t0 = (int)R.term.size();
                        coarse_arrays(lc, d, fv, arrs);
                        for (auto& a : arrs) {
                            int b = h.bcoarse[d][a.second];
                            if (b != 0) R.term.push_back({a.second, b == 1 ? 1 : 0, 0, -1.0});
                        }
                        const int nfc = 1 << (D - 1);
                        for (int c = 0; c < nfc; ++c) {
                            int ff[3] = {2 * fv[0], 2 * fv[1], 2 * fv[2]};
                            int bp = 0;
                            for (int a = 0; a < D; ++a) {
                                if (a == d) continue;
                                ff[a] += (c >> bp) & 1;
                                bp++;
                            }
                            long f = h.face_of(B, d, ff);
                            int b = h.bfine[d][f];
                            if (b != 0) R.term.push_back({f, (b == 1 ? 1 : 0) | (h.cidx[d][f] << 1), 1, 1.0});
                        }
                        const int t1 = (int)R.term.size();
                        if (t1 == t0) return;
                        const int a0 = (int)R.arr.size();
                        for (auto& a : arrs) R.arr.push_back(a.second);
                        R.rec.push_back({d, a0, (int)R.arr.size(), t0, t1});
                    });
                }
        });
        // register terms landing on coarse faces, per direction: (face,
        // moment, coefficient), grouped by face below
        struct Land { long face; int mom; double coef; };
        std::vector<std::vector<Land>> land(D);
        for (const PatchRegs& R : pr)
            for (const Rec& r : R.rec) {
                int ids[8];   // at most 4 coarse face arrays and 4 fine faces
                if (r.t1 - r.t0 > 8) throw std::runtime_error("register face with more than 8 terms");
                for (int t = r.t0; t < r.t1; ++t) {
                    const Term& e = R.term[(size_t)t];
                    ids[t - r.t0] = moment(e.kind, r.d, e.face, e.sel, e.kind ? lf : lc);
                }
                for (int a = r.a0; a < r.a1; ++a)
                    for (int t = r.t0; t < r.t1; ++t)
                        land[r.d].push_back({R.arr[(size_t)a], ids[t - r.t0], R.term[(size_t)t].coef});
            }
        parallel_for(D, [&](int d) {
            std::stable_sort(land[d].begin(), land[d].end(), [](const Land& a, const Land& b) { return a.face < b.face; });
        });
        // the entries of face f of direction d: [first, last)
        auto land_of = [&](int d, long f) {
            auto lo = std::lower_bound(land[d].begin(), land[d].end(), f,
                                       [](const Land& e, long v) { return e.face < v; });
            auto hi = lo;
            while (hi != land[d].end() && hi->face == f) ++hi;
            return std::make_pair(lo, hi);
        };
        // corrected coarse cells: reference L2ProjInterp::reflux applies the
        // correction only where a face of the cell is flagged in its own
        // patch's (ownership-filtered) cf_face_b_coarse (unless
        // conservative_reflux: every cell the registers reach).  Cells covered by
        // the fine level are left out: the average down that follows the
        // reflux (AmrDG.h:1000) overwrites them, so their correction has no
        // effect, and leaving them out lets the reflux run beside it.  Only
        // the valid cells beside a face with register terms can be corrected;
        // they are visited in slot order (patch by patch, x fastest), in
        // chunks on the host threads, concatenated in chunk order.
        // (the slots beside the faces lie on level lc: marked, then read in order)
        const long s0 = h.lev[lc].slot_off, nsl = h.lev[lc].nslot;
        std::vector<uint8_t> hit((size_t)nsl, 0);
        for (int d = 0; d < D; ++d)
            for (size_t k = 0; k < land[d].size(); ++k) {
                if (k > 0 && land[d][k].face == land[d][k - 1].face) continue;
                for (int side = 0; side < 2; ++side) hit[(size_t)(h.face_lr[d][2 * land[d][k].face + side] - s0)] = 1;
            }
        std::vector<int> cand;
        for (long i = 0; i < nsl; ++i)
            if (hit[(size_t)i]) cand.push_back((int)(s0 + i));
        struct Rows { std::vector<int> mom, slot, start, count; std::vector<double> coef; };
        const long ncand = (long)cand.size();
        const int nck = (int)std::min<long>(256, std::max<long>(1, ncand / 2048));
        std::vector<Rows> rows((size_t)nck);
        parallel_for(nck, [&](int k) {
            Rows& O = rows[(size_t)k];
            std::vector<std::pair<int, double>> reg;   // (moment, signed coefficient), summed per moment below
            for (long ci = ncand * k / nck; ci < ncand * (k + 1) / nck; ++ci) {
                const int slot = cand[(size_t)ci];
                int cp = h.lev[lc].p_begin, hi_p = h.lev[lc].p_end - 1;   // patch holding the slot
                while (cp < hi_p) {
                    const int mid = (cp + hi_p + 1) / 2;
                    if (h.patch[mid].off <= slot) cp = mid; else hi_p = mid - 1;
                }
                const PatchLayout& P = h.patch[cp];
                int iv[3] = {0, 0, 0};
                long r = slot - P.off;
                bool valid = true;
                for (int d = 0; d < D; ++d) {
                    iv[d] = P.lo[d] - 1 + (int)(r % P.g[d]);
                    r /= P.g[d];
                    if (iv[d] < P.lo[d] || iv[d] > P.hi[d]) valid = false;
                }
                if (!valid || covered(lc, iv)) continue;
                bool at_if = false;
                reg.clear();
                for (int d = 0; d < D; ++d) {
                    int hv[3] = {iv[0], iv[1], iv[2]};
                    hv[d] += 1;
                    long flo = h.face_of(cp, d, iv), fhi = h.face_of(cp, d, hv);
                    if (h.bcoarse[d][flo] != 0 || h.bcoarse[d][fhi] != 0) at_if = true;
                    auto rl = land_of(d, flo);
                    for (auto e = rl.first; e != rl.second; ++e) reg.push_back({e->mom, e->coef});
                    auto rh = land_of(d, fhi);
                    for (auto e = rh.first; e != rh.second; ++e) reg.push_back({e->mom, -e->coef});
                }
                if (!at_if && !conservative_reflux) continue;
                std::sort(reg.begin(), reg.end(),
                          [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
                const int start = (int)O.mom.size();
                for (size_t k2 = 0; k2 < reg.size();) {
                    size_t e = k2;
                    double sum = 0.0;   // sums of +-1: exact in any order
                    while (e < reg.size() && reg[e].first == reg[k2].first) sum += reg[e++].second;
                    if (sum != 0.0) { O.mom.push_back(reg[k2].first); O.coef.push_back(sum); }
                    k2 = e;
                }
                const int cnt = (int)O.mom.size() - start;
                if (cnt == 0) continue;
                O.slot.push_back((int)h.slot_of(cp, iv));
                O.start.push_back(start);
                O.count.push_back(cnt);
            }
        });
        for (const Rows& O : rows) {
            const int base = (int)h.re_mom.size();
            h.re_mom.insert(h.re_mom.end(), O.mom.begin(), O.mom.end());
            h.re_coef.insert(h.re_coef.end(), O.coef.begin(), O.coef.end());
            for (size_t i = 0; i < O.slot.size(); ++i) {
                h.rc_slot.push_back(O.slot[i]);
                h.rc_lev.push_back((uint8_t)lc);
                h.rc_start.push_back(base + O.start[i]);
                h.rc_count.push_back(O.count[i]);
            }
        }
    }
    for (size_t m = 0; m < h.mom_face.size(); ++m) mom_id[h.mom_kind[m]][h.mom_d[m]][(size_t)h.mom_face[m]] = -1;
    // ---- predictor slots and trace sources (copy ghosts skipped)
    {
        std::vector<int> src_of((size_t)h.NS, -1);
        for (size_t g = 0; g < h.g_slot.size(); ++g)
            if ((h.g_kind[g] & 3) == 0 && (h.g_kind[g] >> 2) == 0) {
                src_of[(size_t)h.g_slot[g]] = h.g_src[g];
                h.copy_ghost.push_back(h.g_slot[g]);
                h.copy_src.push_back(h.g_src[g]);
            }
        h.pred_off.assign((size_t)nlev + 1, 0);
        for (int l = 0; l < nlev; ++l) {
            h.pred_off[(size_t)l] = (long)h.pred_slots.size();
            for (long s = h.lev[l].slot_off; s < h.lev[l].slot_off + h.lev[l].nslot; ++s)
                if (src_of[(size_t)s] < 0) h.pred_slots.push_back((int)s);
        }
        h.pred_off[(size_t)nlev] = (long)h.pred_slots.size();
        for (int d = 0; d < D; ++d) {
            const std::vector<int>& lr = h.face_lr[d];
            std::vector<int>& ft = h.face_tr[d];
            const long nft = (long)lr.size();
            ft.resize((size_t)nft);   // every entry written below
            const int nch = (int)std::min<long>(256, std::max<long>(1, nft / 65536));
            parallel_for(nch, [&](int k) {
                for (long i = nft * k / nch; i < nft * (k + 1) / nch; ++i) {
                    const int s = lr[(size_t)i];
                    ft[(size_t)i] = src_of[(size_t)s] >= 0 ? src_of[(size_t)s] : s;
                }
            });
        }
        // the limiter's cells and neighbours (FillBoundary's copies resolved)
        const LevelLayout& fl = h.lev[nlev - 1];
        const long n = fl.valid_end - fl.valid_begin;
        h.lim_cells.resize((size_t)n);   // every entry written below
        h.lim_nbr.resize((size_t)2 * D * n);
        const int nchunk = (int)std::min<long>(256, std::max<long>(1, n / 4096));
        parallel_for(nchunk, [&](int k) {
            for (long i = n * k / nchunk; i < n * (k + 1) / nchunk; ++i) {
                const long v = fl.valid_begin + i;
                h.lim_cells[(size_t)i] = h.valid_list[(size_t)v];
                for (int d = 0; d < D; ++d)
                    for (int side = 0; side < 2; ++side) {
                        const int f = h.cell_face[d][side][(size_t)v];
                        int sl = h.face_lr[d][2 * (size_t)f + side];
                        if (src_of[(size_t)sl] >= 0) sl = src_of[(size_t)sl];
                        h.lim_nbr[(size_t)(2 * d + side) * n + i] = sl;
                    }
            }
        });
    }
    // centres of the uncovered cells; patches are in slot order, so a slot's
    // patch is the last one whose offset does not exceed it
    {
        const long nu = (long)h.uncovered.size();
        h.unc_centre.resize((size_t)nu * 3);   // every entry written below
        const int nchunk = (int)std::min<long>(256, std::max<long>(1, nu / 4096));
        parallel_for(nchunk, [&](int k) {
            for (long i = nu * k / nchunk; i < nu * (k + 1) / nchunk; ++i) {
                const long s = h.uncovered[(size_t)i];
                auto it = std::upper_bound(h.patch.begin(), h.patch.end(), s,
                                           [](long v, const PatchLayout& pt) { return v < pt.off; });
                const PatchLayout& pt = *(it - 1);
                const long loc = s - pt.off;
                const long r[3] = {loc % pt.g[0], (loc / pt.g[0]) % pt.g[1], loc / ((long)pt.g[0] * pt.g[1])};
                for (int d = 0; d < 3; ++d)
                    h.unc_centre[(size_t)i * 3 + d] =
                        d < D ? h.plo[d] + ((double)(pt.lo[d] + r[d] - 1) + 0.5) * h.lev[pt.lev].dx[d] : 0.0;
            }
        });
    }
    for (int l = 0; l < nlev; ++l) {
        h.grid_slot[l].swap(G[l].slot);
        h.grid_patch[l].swap(G[l].pidx);
    }
    return h;
}

}  // namespace adg
