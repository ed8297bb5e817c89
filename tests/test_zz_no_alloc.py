"""orac_hip_fft_pme_noalloc (static scratch region, no allocation) against
orac_hip_fft_pme on the same inputs: two calls in a row, the second with a
smaller system and grid, each bit-identical to the allocating entry point
(the same kernels on the same fixed-point grid), and sizes beyond the region
rejected with status 2 before any device work."""
import numpy as np
import pytest

from conftest import rel, relrms

pytestmark = pytest.mark.gpu


def _call(ctx, s, noalloc, f0, st0):
    n = s.natoms
    cm = s.molecule_coms()
    fpx, fpy, fpz = (f0[k].copy() for k in range(3))
    st = np.asfortranarray(st0.copy())
    fn = ctx.fft_pme_noalloc if noalloc else ctx.fft_pme
    eer = fn(n, s.xp0, s.yp0, s.zp0, cm[0], cm[1], cm[2], s.charge, s.co, s.oc, s.volume, s.alpha, s.order, *s.nfft,
             fpx, fpy, fpz, st, s.atomp, s.grppt, True, 1e9)
    return eer, np.stack([fpx, fpy, fpz]), st


def test_noalloc_matches_fft_pme_two_calls(gpu_ctx):
    from orac_hip import synth
    big = synth.config("solvated24k")           # 24,460 atoms, 64^3
    small = synth.config("water216")            # 648 atoms, 20^3: the second call is smaller
    rng = np.random.default_rng(7)
    for s in (big, small):
        print("case", s.natoms, "atoms", s.nfft, flush=True)
        gpu_ctx.fft_pme_init(s.natoms, *s.nfft, s.order)
        f0 = rng.standard_normal((3, s.natoms))  # the caller's forces: both entry points add to them
        st0 = rng.standard_normal((3, 3))
        e_ref, f_ref, st_ref = _call(gpu_ctx, s, False, f0, st0)
        e_na, f_na, st_na = _call(gpu_ctx, s, True, f0, st0)
        print("eer", e_ref, e_na, "f relrms", relrms(f_na, f_ref), flush=True)
        assert e_na == e_ref
        assert np.array_equal(f_na, f_ref)
        assert np.array_equal(st_na, st_ref)
        assert rel(e_na, e_ref) == 0.0


def test_noalloc_rejects_sizes_beyond_the_region(gpu_ctx):
    from orac_hip import synth
    from orac_hip.api import OracHipError
    s = synth.config("water216")
    # more atoms than the region takes (checked on the host before any device work)
    gpu_ctx.fft_pme_init(s.natoms, *s.nfft, s.order)
    n = 65537
    z = np.zeros(n)
    one = np.ones(n, np.int32)
    f = [np.zeros(n) for _ in range(3)]
    st = np.zeros((3, 3), order="F")
    with pytest.raises(OracHipError) as e:
        gpu_ctx.fft_pme_noalloc(n, z, z, z, z, z, z, z, s.co, s.oc, s.volume, s.alpha, s.order, *s.nfft, *f, st, one,
                                one, True, 1e9)
    assert e.value.code == 2
    # a grid larger than 64^3 points
    gpu_ctx.fft_pme_init(s.natoms, 72, 72, 72, s.order)
    f = [np.zeros(s.natoms) for _ in range(3)]
    cm = s.molecule_coms()
    with pytest.raises(OracHipError) as e:
        gpu_ctx.fft_pme_noalloc(s.natoms, s.xp0, s.yp0, s.zp0, cm[0], cm[1], cm[2], s.charge, s.co, s.oc, s.volume,
                                s.alpha, s.order, 72, 72, 72, *f, st, s.atomp, s.grppt, True, 1e9)
    assert e.value.code == 2
