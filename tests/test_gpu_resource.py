"""Plans with rev1 resource files on the GPU (src/wavemoth.c:556-569):
- a file given, or found under the configured directory, is read and
  validated and leaves the synthesis unchanged (the device path generates
  Lambda itself);
- WAVEMOTH_RESOURCE_APPLY runs the Legendre stage from the file's strips
  (the reference's pull_a_through_legendre_block, src/wavemoth.c:958-1009):
  its work buffer against the oracle's Legendre stage and its map against
  the oracle and the on-the-fly FP64 path, to 1e-12 rel rms;
- with butterfly-compressed matrices (src/butterfly.c.in:284-432) the
  device applies the trees: work against the independent numpy walk of the
  same file to 1e-12 and against the oracle to the compression precision
  (1e-9 at ID precision 1e-12), map against the on-the-fly path;
- the flag's preconditions are refused with a message."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

APPLY = 0x400


def _file(tmp_path, N, L, M):
    from wavemoth_amd.resource import write_resource
    fn = str(tmp_path / ("%d_%d_%d.dat" % (N, L, M)))
    write_resource(fn, N, L, M)
    return fn


@pytest.mark.parametrize("N,L,M,K", [(16, 40, 40, 3), (32, 64, 50, 1), (64, 128, 128, 2)])
def test_resource_apply_matches_oracle(tmp_path, oracle, N, L, M, K):
    import wavemoth_amd as wm
    fn = _file(tmp_path, N, L, M)
    alm = wm.power_law_alm(L, M, K, 300 + N)
    print("case resource apply N=%d L=%d M=%d K=%d" % (N, L, M, K), flush=True)
    out = np.zeros((12 * N * N, K))
    p = wm.ShtPlan(N, L, M, alm, out, 'mmajor', matrix_data_filename=fn, flags=APPLY)
    got = p.execute().copy()
    work = p.get_work()
    op = oracle.OraclePlan(N, L, M, K)
    ref_work = op.legendre_stage(alm)
    ref = op.synthesize(alm)
    ew = np.linalg.norm(work - ref_work) / np.linalg.norm(ref_work)
    em = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    out2 = np.zeros_like(out)
    fly = wm.ShtPlan(N, L, M, alm, out2, 'mmajor').execute()
    ef = np.linalg.norm(got - fly) / np.linalg.norm(fly)
    print("work %.3e  map vs oracle %.3e  vs on-the-fly path %.3e" % (ew, em, ef), flush=True)
    assert ew <= 1e-12 and em <= 1e-12 and ef <= 1e-12


def test_resource_given_or_configured_leaves_map_unchanged(tmp_path):
    import wavemoth_amd as wm
    from wavemoth_amd import _native
    from wavemoth_amd.resource import resource_default_path
    N, L, K = 16, 32, 2
    fn = _file(tmp_path, N, L, L)
    alm = wm.power_law_alm(L, L, K, 9)
    plain = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor').execute().copy()
    given = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor',
                       matrix_data_filename=fn).execute().copy()
    assert np.array_equal(given, plain)
    # Nside < 0 takes the file's (src/wavemoth.c:561-562)
    lib = _native.lib()
    out = np.zeros((12 * N * N, K))
    p = lib.wavemoth_plan_to_healpix(-1, L, L, K, 1, alm.ctypes.data, out.ctypes.data, 0, 0,
                                     fn.encode())
    assert p, _native.last_error()
    assert lib.wavemoth_plan_info(p, 0) == N
    lib.wavemoth_execute(p)
    lib.wavemoth_destroy_plan(p)
    assert np.array_equal(out, plain)
    # found under the configured directory
    d = tmp_path / "res"
    os.makedirs(d / "rev1")
    os.replace(fn, resource_default_path(str(d), N))
    try:
        lib.wavemoth_configure(str(d).encode())
        out = np.zeros((12 * N * N, K))
        p = lib.wavemoth_plan_to_healpix(N, L, L, K, 1, alm.ctypes.data, out.ctypes.data, 0,
                                         APPLY, None)
        assert p, _native.last_error()
        lib.wavemoth_execute(p)
        lib.wavemoth_destroy_plan(p)
        assert np.linalg.norm(out - plain) <= 1e-12 * np.linalg.norm(plain)
    finally:
        lib.wavemoth_configure(b"")


def test_resource_apply_preconditions(tmp_path):
    from wavemoth_amd import _native
    lib = _native.lib()
    assert not lib.wavemoth_plan_to_healpix(8, 16, 16, 1, 1, None, None, 0, APPLY, None)
    assert "needs a resource file" in _native.last_error()
    fn = _file(tmp_path, 8, 16, 16)
    assert not lib.wavemoth_plan_to_healpix(8, 16, 16, 16, 1, None, None, 0, APPLY | 0x1000,
                                            fn.encode())
    assert "no WAVEMOTH_MIXED" in _native.last_error()


@pytest.mark.parametrize("N,L,M,K,cs,levels", [
    (8, 16, 16, 2, 2, 1),
    (16, 32, 32, 3, 4, 2),
    (16, 32, 24, 1, 4, 3),      # full depth: leaves under every root
    (32, 64, 64, 2, 8, 2),      # forest: 2 roots per matrix
    (64, 128, 100, 9, 16, 2),   # recurrence strips in the remainders; 9 maps (two map blocks)
])
def test_compressed_resource_apply(tmp_path, oracle, N, L, M, K, cs, levels):
    import wavemoth_amd as wm
    from wavemoth_amd.resource import write_resource
    from test_resource_tree import _work_from_tree_file
    fn = str(tmp_path / "c.dat")
    write_resource(fn, N, L, M, chunk_size=cs, eps=1e-12, num_levels=levels)
    alm = wm.power_law_alm(L, M, K, 40 + N)
    print("case compressed apply N=%d L=%d M=%d K=%d chunk=%d levels=%d" % (N, L, M, K, cs,
                                                                           levels), flush=True)
    out = np.zeros((12 * N * N, K))
    p = wm.ShtPlan(N, L, M, alm, out, 'mmajor', matrix_data_filename=fn, flags=APPLY)
    got = p.execute().copy()
    work = p.get_work()
    tree_work, _ = _work_from_tree_file(fn, alm)
    ref_work = oracle.OraclePlan(N, L, M, K).legendre_stage(alm)
    et = np.linalg.norm(work - tree_work) / np.linalg.norm(tree_work)
    ew = np.linalg.norm(work - ref_work) / np.linalg.norm(ref_work)
    fly = wm.ShtPlan(N, L, M, alm, np.zeros_like(out), 'mmajor').execute()
    ef = np.linalg.norm(got - fly) / np.linalg.norm(fly)
    print("work vs numpy tree walk %.3e  vs oracle %.3e  map vs on-the-fly %.3e" % (et, ew, ef),
          flush=True)
    assert et <= 1e-12 and ew <= 1e-9 and ef <= 1e-9
