"""The allocation-free multispecies dSdt (ocfd_compute_dsdt_multi_noalloc):
the same results as ocfd_compute_dsdt_multi bit for bit, from two calls in a
row of which the second has the smaller input (the static scratch keeps the
first call's contents), on regular and cut boxes, inviscid and viscous; and a
level whose scratch need exceeds the static array is refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _chem_level(case, eb=None):
    import ocfd
    boxes, g, kind = case[:3]
    lev = ocfd.Level(boxes, g, eb=eb)
    p = ocfd.default_parm()
    S, P = lev.new_mf(), lev.new_spec_mf()
    lev.init_species_state(kind, S, P, p)
    lev.fill_boundary(S, 0.0)
    lev.fill_boundary_spec(P, 0.0)
    return lev, S, P, p


def _dsdt(lev, S, P, p, flags, noalloc):
    import torch
    import ocfd
    D, Ds = lev.new_mf(ngrow=0), lev.new_mf(ngrow=0, ncomp=ocfd.NSPEC)
    fn = ocfd.lib().ocfd_compute_dsdt_multi_noalloc if noalloc else ocfd.lib().ocfd_compute_dsdt_multi
    ocfd.check(fn(lev.handle, ocfd._ptr(S), ocfd._ptr(P), ocfd._ptr(D), ocfd._ptr(Ds), 0.0, ctypes.byref(p), flags,
                  None), "compute_dsdt_multi" + ("_noalloc" if noalloc else ""))
    torch.cuda.synchronize()
    return D.cpu().numpy(), Ds.cpu().numpy()


@pytest.mark.parametrize("visc", [0, 1])
def test_noalloc_matches_allocating_entry_point(torch_cuda, visc):
    import ocfd
    from ocfd import cases
    flags = ocfd.VISC if visc else 0
    big = _chem_level(cases.chemsod(256, 8, 8, max_grid=64))
    small = _chem_level(cases.chemsod(128, 4, 4, max_grid=32))
    ref_big = _dsdt(*big, flags, noalloc=False)
    ref_small = _dsdt(*small, flags, noalloc=False)
    got_big = _dsdt(*big, flags, noalloc=True)        # two calls in a row,
    got_small = _dsdt(*small, flags, noalloc=True)    # the second on the smaller level
    for got, ref in ((got_big, ref_big), (got_small, ref_small)):
        for a, b in zip(got, ref):
            assert np.isfinite(b).all()
            assert np.array_equal(a, b)


def test_noalloc_on_cut_boxes(torch_cuda):
    """An embedded-boundary level: its cut boxes run the EB-CHEM kernels,
    which use the same scratch."""
    import ocfd
    from ocfd import cases
    boxes, _, _, sp = cases.eb_channel(64, 16, 2, max_grid=16)
    # the channel with outflow at both x ends, so the species data of the
    # shock-tube initialisation stay consistent in the ghost cells
    g = ocfd.make_geom((64, 16, 2), (0.0, 0.0, 0.0), (2.0, 0.5, 2 * 2.0 / 64), periodic=(0, 0, 1),
                       lo_bc=(cases.OUTFLOW, cases.SLIPWALL, cases.INTERIOR),
                       hi_bc=(cases.OUTFLOW, cases.SLIPWALL, cases.INTERIOR))
    lev, S, P, p = _chem_level((boxes, g, ocfd.CASE_CHEMSOD), eb=sp)
    assert 1 in lev.box_types()
    ref = _dsdt(lev, S, P, p, ocfd.VISC, noalloc=False)
    got = _dsdt(lev, S, P, p, ocfd.VISC, noalloc=True)
    for a, b in zip(got, ref):
        assert np.isfinite(b).all()
        assert np.array_equal(a, b)


def test_noalloc_refuses_a_level_beyond_its_scratch(torch_cuda):
    import ocfd
    from ocfd import cases
    lev, S, P, p = _chem_level(cases.chemsod(1280, 16, 8, max_grid=128))
    assert lev.mf_size() // ocfd.NCONS * 488 > 64 << 20
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        _dsdt(lev, S, P, p, 0, noalloc=True)
