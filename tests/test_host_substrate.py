"""CPU checks of the C-ABI library: it loads, exports every entry point its
headers declare, and its host AMR substrate (adg_set_levels lists) builds
the same coarse-fine face masks as the oracle's restatement of
AMR_set_flux_registers (AmrDG_Level.tpp:240-414).  No device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from layouts import kh_bands, single_level, tiles, two_level_centre

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = []
    for h in ("adg.h", "adg_host.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        syms += re.findall(r"\b(adg_[A-Za-z_0-9]+)\s*\(", txt)
    return sorted(set(syms))


def test_library_exports_header_symbols(adg):
    lib = C.CDLL(adg.LIBPATH)
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert set(adg.EXPORTS) <= set(syms)


def inspect(adg, D, dom, levels, lev, d, kind, periodic=None):
    lib = adg.lib()
    fn = lib.adg_host_inspect
    fn.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                   C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
    fn.restype = C.c_int
    npatch = np.array([len(l) for l in levels], dtype=np.int32)
    flat = np.array([v for l in levels for lo, hi in l for v in list(lo) + list(hi)], dtype=np.int32)
    per = np.array(periodic or [1] * D, dtype=np.int32)
    dom = np.array(dom, dtype=np.int32)
    out = np.zeros(1 << 20, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    n = fn(D, ip(dom), ip(per), len(levels), ip(npatch), ip(flat), lev, d, kind, ip(out), out.size)
    assert n >= 0
    return out[:n].copy()


LAYOUTS = {
    "two_level": (two_level_centre(8, 4), [8, 8]),
    "kh_bands": (kh_bands(16, 3, 8), [16, 16]),
    "corner_patch": ([tiles(8, 4), [((0, 0), (5, 7)), ((10, 2), (15, 13))]], [8, 8]),
    "wrap_edge": ([tiles(8, 8), [((12, 4), (15, 9))]], [8, 8]),
}


@pytest.mark.parametrize("periodic", [[1, 1], [0, 0], [1, 0]])
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_face_masks_match_oracle(adg, orc, name, periodic):
    levels, dom = LAYOUTS[name]
    H = orc.Hier(2, levels, dom, [0, 0], [1, 1], periodic=periodic)
    for l in range(len(levels)):
        for d in range(2):
            for kind in range(3):
                got = inspect(adg, 2, dom, levels, l, d, kind, periodic=periodic)
                ref = H.face_mask(l, d, kind)
                assert np.array_equal(got, ref), (name, l, d, kind)


def test_ownership_rule_skips_patch_lo_face(adg):
    """cf_face_b_coarse never flags a coarse patch's own lo face
    (AmrDG_Level.tpp:330-343): with coarse patches split at x=4 and a fine
    patch right of x=4, the C-F face at x=4 is flagged only in the left
    patch's array (as its hi face)."""
    levels = [[((0, 0), (3, 7)), ((4, 0), (7, 7))], [((8, 4), (11, 9))]]
    left = inspect(adg, 2, [8, 8], levels, 0, 0, 0)
    nf_left = 5 * 8
    left_patch = left[:nf_left].reshape(8, 5)    # [j][i], faces i = 0..4
    right_patch = left[nf_left:].reshape(8, 5)   # [j][i], faces i = 4..8
    assert (left_patch[2:5, 4] == 1).all()       # coarse cells 2..4 in y see fine to their right
    assert (right_patch[:, 0] == 0).all()        # same faces as the right patch's lo face: skipped
    assert (right_patch[2:5, 2] == -1).all()     # fine region ends at x=6: flagged inside the patch


def test_counts(adg):
    levels = kh_bands(16, 3, 8)
    c = inspect(adg, 2, [16, 16], levels, 0, 0, 3)
    ns, ng = c[0], c[1]
    nvalid = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) for l in levels for lo, hi in l)
    assert ns == nvalid + ng
    assert c[5] > 0 and c[6] > 0 and c[7] < nvalid


def test_bad_layouts_rejected(adg):
    with pytest.raises(AssertionError):
        inspect(adg, 2, [8, 8], [tiles(8, 8), [((3, 2), (6, 5))]], 0, 0, 3)   # fine box not ratio-aligned
    with pytest.raises(AssertionError):
        inspect(adg, 2, [8, 8], [[((0, 0), (3, 7))], []], 0, 0, 3)            # level 0 incomplete -> ghosts fail


TABLES = ["nodes", "weights", "V", "Vinv", "Mk_h_w", "Mk_h_w_inv", "Mk_pred", "Sk_predVinv", "Mk_corr", "Sk_corr",
          "Mkbdm", "Mkbdp", "quadmat", "Minv", "P_cf", "P_fc", "P_flux_low", "P_flux_high"]


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("P", [1, 2, 3])
def test_product_tables_match_oracle(adg, orc, P, D):
    """The element tables the library builds (host side of adg_create) equal
    the oracle's restatement of AmrDG_ElementMatrix.tpp / AmrDG_Interpolater.tpp."""
    lib = adg.lib()
    fn = lib.adg_host_table
    fn.argtypes = [C.c_int, C.c_int, C.c_char_p, C.POINTER(C.c_double), C.c_int]
    fn.restype = C.c_int
    T = orc.Tables(P, D)
    for name in TABLES:
        ref = T.get(name)
        n = fn(P, D, name.encode(), None, 0)
        assert n == ref.size, (name, n, ref.size)
        got = np.zeros(n)
        fn(P, D, name.encode(), got.ctypes.data_as(C.POINTER(C.c_double)), n)
        scale = max(np.abs(ref).max(), 1.0)
        assert np.abs(got - ref.ravel()).max() <= 1e-12 * scale, name
