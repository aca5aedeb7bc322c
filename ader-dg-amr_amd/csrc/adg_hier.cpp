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
                          const int* per, int nlev, const int* npatch, const int* boxes, int independent, int conservative_reflux) {
    if (D < 2 || D > 3 || nlev < 1 || nlev > 8) throw std::runtime_error("bad dim or level count");
    if (independent && nlev != 1) throw std::runtime_error("independent patches form a single level");
    Hierarchy h;
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
    for (int l = 0; l < nlev; ++l) {
        Grid& g = G[l];
        g.D = D;
        long nd = 1;
        for (int d = 0; d < 3; ++d) { g.dom[d] = d < D ? h.lev[l].dom[d] : 1; nd *= g.dom[d]; }
        g.slot.assign(nd, -1);
        g.pidx.assign(nd, -1);
        for (int p = h.lev[l].p_begin; p < h.lev[l].p_end; ++p)
            for_box(D, h.patch[p].lo, h.patch[p].hi, [&](const int* iv) {
                long c = g.lin(iv);
                if (g.slot[c] >= 0) throw std::runtime_error("overlapping boxes on a level");
                g.slot[c] = (int)h.slot_of(p, iv);
                g.pidx[c] = p;
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
    h.slot_lev.assign(h.NS, 0);
    for (int l = 0; l < nlev; ++l)
        std::fill(h.slot_lev.begin() + h.lev[l].slot_off, h.slot_lev.begin() + h.lev[l].slot_off + h.lev[l].nslot,
                  (uint8_t)l);
    for (int d = 0; d < D; ++d)
        for (int s = 0; s < 2; ++s) h.cell_face[d][s].clear();
    for (int l = 0; l < nlev; ++l) {
        h.lev[l].valid_begin = (long)h.valid_list.size();
        for (int p = h.lev[l].p_begin; p < h.lev[l].p_end; ++p)
            for_box(D, h.patch[p].lo, h.patch[p].hi, [&](const int* iv) {
                h.valid_list.push_back((int)h.slot_of(p, iv));
                for (int d = 0; d < D; ++d) {
                    int hv[3] = {iv[0], iv[1], iv[2]};
                    hv[d] += 1;
                    h.cell_face[d][0].push_back((int)h.face_of(p, d, iv));
                    h.cell_face[d][1].push_back((int)h.face_of(p, d, hv));
                }
                if (!covered(l, iv)) h.uncovered.push_back((int)h.slot_of(p, iv));
            });
        h.lev[l].valid_end = (long)h.valid_list.size();
    }
    // ---- faces
    for (int d = 0; d < D; ++d) {
        h.face_lr[d].assign(2 * h.NF[d], 0);
        h.face_lev[d].assign(h.NF[d], 0);
        for (int p = 0; p < (int)h.patch.size(); ++p) {
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
        }
    }
    // ---- ghost sources (FillPatchSingleLevel / FillPatchTwoLevels semantics)
    for (int p = 0; p < (int)h.patch.size(); ++p) {
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
                h.g_slot.push_back((int)s); h.g_src.push_back((int)h.slot_of(p, w));
                h.g_kind.push_back(0); h.g_child.push_back(0);
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
                h.g_slot.push_back((int)s); h.g_src.push_back((int)src);
                h.g_kind.push_back((uint8_t)(0 | (mask << 2))); h.g_child.push_back(0);
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
            h.g_slot.push_back((int)s); h.g_src.push_back((int)cs);
            h.g_kind.push_back((uint8_t)(1 | (mask << 2))); h.g_child.push_back((uint8_t)child);
        });
    }
    // ---- average down groups
    const int nch = 1 << D;
    h.ad_coarse.assign(nlev, {});
    h.ad_fine.assign(nlev, {});
    for (int l = 1; l < nlev; ++l)
        for (int p = h.lev[l].p_begin; p < h.lev[l].p_end; ++p) {
            int clo[3], chi[3];
            for (int d = 0; d < D; ++d) { clo[d] = h.patch[p].lo[d] >> 1; chi[d] = h.patch[p].hi[d] >> 1; }
            for_box(D, clo, chi, [&](const int* iv) {
                h.ad_coarse[l].push_back((int)find(l - 1, iv));
                for (int c = 0; c < nch; ++c) {
                    int fv[3] = {2 * iv[0], 2 * iv[1], 2 * iv[2]};
                    for (int d = 0; d < D; ++d) fv[d] += (c >> (D - 1 - d)) & 1;
                    h.ad_fine[l].push_back((int)h.slot_of(p, fv));
                }
            });
        }
    // ---- C-F face masks (AmrDG_Level.tpp:240-414)
    for (int d = 0; d < D; ++d) {
        h.bcoarse[d].assign(h.NF[d], 0);
        h.bfine[d].assign(h.NF[d], 0);
        h.cidx[d].assign(h.NF[d], 0);
    }
    for (int p = 0; p < (int)h.patch.size(); ++p) {
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
    }
    // ---- flux-register algebra (FluxRegister CrseAdd(-1) / FineAdd(+1) /
    // Reflux(vol=1) as called in AmrDG.h:1138-1153 and AmrDG_Level.tpp:416-446)
    std::map<std::tuple<int, int, long>, int> mom_id;  // (kind, d, face) -> index
    auto moment = [&](int kind, int d, long face, int sel, int lev) {
        auto key = std::make_tuple(kind, d, face);
        auto it = mom_id.find(key);
        if (it != mom_id.end()) return it->second;
        int id = (int)h.mom_face.size();
        mom_id[key] = id;
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
        for (int ii = 0; ii < nimg; ++ii) {
            int fv[3] = {fv0[0], fv0[1], fv0[2]};
            fv[d] = imgs[ii];
            for (int cp = h.lev[lc].p_begin; cp < h.lev[lc].p_end; ++cp) {
                const PatchLayout& C = h.patch[cp];
                bool in = true;
                for (int a = 0; a < D; ++a)
                    if (fv[a] < C.lo[a] || fv[a] > C.hi[a] + (a == d ? 1 : 0)) in = false;
                if (in) out.push_back({cp, h.face_of(cp, d, fv)});
            }
        }
    };
    for (int lf = 1; lf < nlev; ++lf) {
        const int lc = lf - 1;
        // per coarse face (direction d): accumulated register terms
        std::vector<std::unordered_map<long, std::vector<std::pair<int, double>>>> land(D);
        std::vector<std::pair<int, long>> arrs;
        for (int B = h.lev[lf].p_begin; B < h.lev[lf].p_end; ++B) {
            int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { clo[d] = h.patch[B].lo[d] >> 1; chi[d] = h.patch[B].hi[d] >> 1; }
            for (int d = 0; d < D; ++d)
                for (int side = 0; side < 2; ++side) {
                    int tlo[3] = {clo[0], clo[1], clo[2]}, thi[3] = {chi[0], chi[1], chi[2]};
                    tlo[d] = thi[d] = side == 0 ? clo[d] : chi[d] + 1;
                    for_box(D, tlo, thi, [&](const int* fv) {
                        std::vector<std::pair<int, double>> reg;
                        coarse_arrays(lc, d, fv, arrs);
                        for (auto& a : arrs) {
                            int b = h.bcoarse[d][a.second];
                            if (b != 0) reg.push_back({moment(0, d, a.second, b == 1 ? 1 : 0, lc), -1.0});
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
                            if (b != 0)
                                reg.push_back({moment(1, d, f, (b == 1 ? 1 : 0) | (h.cidx[d][f] << 1), lf), 1.0});
                        }
                        if (reg.empty()) return;
                        for (auto& a : arrs) {
                            auto& v = land[d][a.second];
                            v.insert(v.end(), reg.begin(), reg.end());
                        }
                    });
                }
        }
        // corrected coarse cells: reference L2ProjInterp::reflux applies the
        // correction only where a face of the cell is flagged in its own
        // patch's (ownership-filtered) cf_face_b_coarse (unless
        // conservative_reflux: every cell the registers reach).  Cells covered by
        // the fine level are left out: the average down that follows the
        // reflux (AmrDG.h:1000) overwrites them, so their correction has no
        // effect, and leaving them out lets the reflux run beside it.
        for (int cp = h.lev[lc].p_begin; cp < h.lev[lc].p_end; ++cp)
            for_box(D, h.patch[cp].lo, h.patch[cp].hi, [&](const int* iv) {
                if (covered(lc, iv)) return;
                bool at_if = false;
                std::map<int, double> acc;
                for (int d = 0; d < D; ++d) {
                    int hv[3] = {iv[0], iv[1], iv[2]};
                    hv[d] += 1;
                    long flo = h.face_of(cp, d, iv), fhi = h.face_of(cp, d, hv);
                    if (h.bcoarse[d][flo] != 0 || h.bcoarse[d][fhi] != 0) at_if = true;
                    auto itl = land[d].find(flo);
                    if (itl != land[d].end())
                        for (auto& e : itl->second) acc[e.first] += e.second;
                    auto ith = land[d].find(fhi);
                    if (ith != land[d].end())
                        for (auto& e : ith->second) acc[e.first] -= e.second;
                }
                if (!at_if && !conservative_reflux) return;
                int start = (int)h.re_mom.size();
                for (auto& e : acc)
                    if (e.second != 0.0) { h.re_mom.push_back(e.first); h.re_coef.push_back(e.second); }
                int cnt = (int)h.re_mom.size() - start;
                if (cnt == 0) return;
                h.rc_slot.push_back((int)h.slot_of(cp, iv));
                h.rc_lev.push_back((uint8_t)lc);
                h.rc_start.push_back(start);
                h.rc_count.push_back(cnt);
            });
    }
    return h;
}

}  // namespace adg
