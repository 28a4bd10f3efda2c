"""The rest of the Nbnxm::gpu_* API through the C ABI: PME load balancing's
parameter update, the non-blocking finish, event timings, the nbat force buffer
of gpu_get_f, the state after a new atom count, and the j > i rule on the
diagonal of a hand-built list."""
import ctypes
import os
import time

import numpy as np
import pytest

import nbnxm_amd as nb
import oracle
from conftest import announce
from helpers import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water3k():
    s = nb.systems.water_box(1000, None, seed=51)
    return s


def _fetch_fd(task):
    f, fd, elj, eel, fshift = task.fetch()
    return fd.reshape(-1, 3), elj, eel, fshift


def test_pme_loadbal_update_param(water3k):
    """gpu_pme_loadbal_update_param (nbnxm_gpu_data_mgmt.cpp:511-527): PME
    tuning raises rcoulomb (new beta at the same tolerance) while rvdw stays,
    which selects the twin cut-off; then the tabulated correction.  Each update
    matches the FP64 oracle at the new constants and equals, bit for bit, a
    fresh task created with them on the same list."""
    s = water3k
    rlist = 1.3
    search, plist, atoms = nb.build(s, rlist, nb.VDW_CUT_COMB_GEOM)
    real = atoms.atom_indices >= 0
    ic0 = nb.interaction_const(1.0, 1e-5)
    task = nb.GpuTask(ic0, atoms, rlist, None, nb.ACCUM_FIXED64)
    task.init_pairlist(plist)
    task.launch()
    f0, _, _, _ = _fetch_fd(task)
    for name, ic in (("twin", nb.interaction_const_ex(1.2, rvdw=1.0, ewald_rtol=1e-5)),
                     ("tabulated twin", nb.interaction_const_ex(1.2, rvdw=1.0, ewald_rtol=1e-5, tabulated=True)),
                     ("back", nb.interaction_const(1.0, 1e-5))):
        announce(f"pme loadbal update -> {name}")
        assert ic.ewaldcoeff_q != ic0.ewaldcoeff_q or name == "back"
        task.pme_loadbal_update_param(ic, rlist_outer=rlist)
        task.launch()
        f, elj, eel, fshift = _fetch_fd(task)
        fresh = nb.GpuTask(ic, atoms, rlist, None, nb.ACCUM_FIXED64)
        fresh.init_pairlist(plist)
        fresh.launch()
        ff, eljf, eelf, fshiftf = _fetch_fd(fresh)
        assert np.array_equal(f.view(np.uint64), ff.view(np.uint64)), name
        assert elj == eljf and eel == eelf and np.array_equal(fshift, fshiftf)
        f64, _, elj64, eel64 = oracle.kernel(plist, atoms, ic, True, "f64")
        assert rel_rms(f[real], f64[real]) < 5e-6
        assert elj == pytest.approx(elj64, rel=2e-5)
        assert eel == pytest.approx(eel64, rel=2e-5)
        expect_ana = 0 if ic.tabulatedEwald else 1
        assert task._lib.nbnxm_gpu_is_kernel_ewald_analytical(task._h) == expect_ana
        del fresh
    # Back at the first constants: the first forces again.
    assert np.array_equal(f.view(np.uint64), f0.view(np.uint64))
    # The LJ-PME rule is fixed at gpu_init; the list radius must cover rc.
    with pytest.raises(nb.NbnxmError, match="LJ-PME"):
        task.pme_loadbal_update_param(nb.interaction_const_ex(1.0, lj_pme=True))
    with pytest.raises(nb.NbnxmError, match="rlistInner"):
        task.pme_loadbal_update_param(nb.interaction_const(1.4, 1e-5), rlist_outer=rlist)


def test_try_finish_task(water3k):
    """gpu_try_finish_task in Check mode (gpu_common.h:253-333): 0 while queued
    work runs, then 1 with the energies and shift forces wait_finish_task gives."""
    announce("try_finish_task water 96k")
    s = nb.systems.config_water_96k()
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    ref = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, None, nb.ACCUM_FP64)
    ref.init_pairlist(plist)
    lib = nb.load()
    ref.launch(compute_energy=True, compute_virial=True)
    nb._lib.check(lib.nbnxm_gpu_launch_cpyback(ref._h, None, None), "cpyback")
    ref_out = ref.wait()
    task = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, None, nb.ACCUM_FP64)
    task.init_pairlist(plist)
    # About 20 FP64 launches (about a millisecond each) are queued before the check.
    for _ in range(20):
        task.clear_outputs()
        task.launch(compute_energy=True, compute_virial=True)
    nb._lib.check(lib.nbnxm_gpu_launch_cpyback(task._h, None, None), "cpyback")
    assert task.try_finish() is None
    deadline = time.time() + 30.0
    out = None
    while out is None and time.time() < deadline:
        time.sleep(0.001)
        out = task.try_finish()
    assert out is not None
    assert out[0] == ref_out[0] and out[1] == ref_out[1]
    assert np.array_equal(out[2], ref_out[2])
    # A launch after a "not ready" check still reports its own status correctly.
    task.launch(compute_energy=False, compute_virial=False)
    task.wait()


def test_timings(water3k):
    """gpu_get_timings / gpu_reset_timings (nbnxm_gpu_data_mgmt.cpp:809-821):
    counts per kernel kind as gpu_accumulate_timings keeps them
    (gpu_common.h:186-250), positive times, off by default."""
    announce("timings water3k")
    s = water3k
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    task = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, 1.05, nb.ACCUM_FIXED64, rolling_parts=2)
    assert task.timings() is None
    task.set_timing(True)
    task.init_pairlist(plist)
    task.copy_xq(atoms.xq)
    task.launch(compute_energy=True, compute_virial=True)  # fresh list: prune + F+E
    task.fetch()
    t = task.timings()
    assert t.pruneTime.c == 1 and t.pruneTime.t > 0
    assert t.ktime[1][1].c == 1 and t.ktime[1][1].t > 0
    assert t.ktime[0][0].c == 0 and t.nb_c == 1
    assert t.pl_h2d_c == 1 and t.pl_h2d_t > 0 and t.nb_h2d_t > 0 and t.nb_d2h_t > 0
    task.prune(2)  # rolling
    task.launch(compute_energy=False, compute_virial=False)
    task.fetch()
    t = task.timings()
    assert t.dynamicPruneTime.c == 1 and t.ktime[0][0].c == 1 and t.nb_c == 2
    assert t.pruneTime.c == 1
    task.reset_timings()
    t = task.timings()
    assert t.nb_c == 0 and t.ktime[1][1].c == 0 and t.ktime[1][1].t == 0.0
    task.set_timing(False)
    assert task.timings() is None


def _hip():
    # The HIP runtime the library is bound to (same soname, already loaded).
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    return hip


def _read_device(ptr, n):
    hip = _hip()
    out = np.zeros(n, dtype=np.float32)
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), 4 * n, 2) == 0  # device to host
    return out


@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT])
def test_get_f(water3k, accumulation):
    """gpu_get_f (gpu_data_mgmt.h:133): the nbat force buffer a
    GpuForceReduction registers holds the launch's forces as floats in grid
    order, accumulates over launches until gpu_clear_outputs as the
    reference's buffer does, and cpyback reads the same values."""
    announce(f"get_f water3k accumulation={accumulation}")
    s = water3k
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    n3 = 3 * atoms.num_atoms
    ref = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, None, accumulation)
    ref.init_pairlist(plist)
    ref.launch(compute_energy=True, compute_virial=True)
    f_ref, _, elj_ref, eel_ref, _ = ref.fetch(want_double=False)
    vir_ref = ref.virial()

    task = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, None, accumulation)
    task.init_pairlist(plist)
    d_f = task.get_f()
    assert d_f and task.get_f() == d_f
    task.clear_outputs()
    task.launch(compute_energy=True, compute_virial=True)
    g = _read_device(d_f, n3)
    f, _, elj, eel, _ = task.fetch(want_double=False)
    tol = 0.0 if accumulation == nb.ACCUM_FIXED64 else 1e-5 * np.abs(f_ref).max()
    assert np.abs(g - f_ref).max() <= tol
    assert np.abs(f - f_ref).max() <= tol
    if accumulation == nb.ACCUM_FIXED64:
        assert elj == elj_ref and eel == eel_ref
        assert np.array_equal(task.virial(), vir_ref)
    # No clear in between: the buffer holds the sum of two launches (the same
    # F+E kernel: its Ewald form differs from the force-only kernel's in the
    # last bits).
    task.launch(compute_energy=True, compute_virial=False)
    g2 = _read_device(d_f, n3)
    assert np.abs(g2 - 2 * f_ref).max() <= 2 * tol + 1e-6 * np.abs(f_ref).max()
    task.clear_outputs()
    task.launch(compute_energy=True, compute_virial=False)
    g3 = _read_device(d_f, n3)
    assert np.abs(g3 - f_ref).max() <= tol
    stream = task.stream()
    assert stream


def test_new_atom_count_needs_new_list(water3k):
    """After gpu_init_atomdata with another grid size the old list and
    buffer-op maps no longer apply: launch, prune and the buffer ops return
    NBNXM_ERR_STATE until gpu_init_pairlist / gpu_init_x_to_nbat_x run again."""
    announce("state after init_atomdata")
    s = water3k
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    task = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms, 1.1, None, nb.ACCUM_FIXED64)
    task.init_pairlist(plist)
    task.init_x_to_nbat_x(s.n, atoms.atom_indices)
    task.launch()
    f_ref, _, _, _ = _fetch_fd(task)
    small = nb.systems.water_box(300, None, seed=52)
    _, plist2, atoms2 = nb.build(small, 1.1, nb.VDW_CUT_COMB_GEOM)
    assert atoms2.num_atoms != atoms.num_atoms
    lib = nb.load()
    nb._lib.check(lib.nbnxm_gpu_init_atomdata(task._h, ctypes.byref(atoms2.view())), "init_atomdata")
    assert lib.nbnxm_gpu_launch_kernel(task._h, 1, 1) == -3
    assert lib.nbnxm_gpu_launch_kernel_pruneonly(task._h, 1) == -3
    assert lib.nbnxm_gpu_x_to_nbat_x(task._h, ctypes.c_void_p(1)) == -3
    assert lib.nbnxm_gpu_add_nbat_f_to_f(task._h, ctypes.c_void_p(1), None, 0) == -3
    # The new box's shift vectors and list make the task usable again.
    nb._lib.check(lib.nbnxm_gpu_upload_shiftvec(task._h, atoms2.shift_vec.ctypes.data), "upload_shiftvec")
    task.atoms = atoms2
    task.init_pairlist(plist2)
    task.launch()
    f2, _, _, _ = _fetch_fd(task)
    fresh = nb.GpuTask(nb.interaction_const(1.0, 1e-5), atoms2, 1.1, None, nb.ACCUM_FIXED64)
    fresh.init_pairlist(plist2)
    fresh.launch()
    ff, _, _, _ = _fetch_fd(fresh)
    assert np.array_equal(f2, ff)


def test_diagonal_without_newton_mask(water3k):
    """A hand-built list whose diagonal groups carry no exclusion mask
    (excl_ind 0): the kernels still evaluate only j > i on the diagonal of the
    central shift (nonSelfInteraction, nbnxm_cuda_kernel.cuh:423,508), as the
    oracle does, instead of counting j <= i pairs and the i = i self pair."""
    s = water3k
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    hand = plist.copy()
    changed = 0
    for e in hand.sci:
        if e["shift"] != nb.CENTRAL_SHIFT:
            continue
        for g in range(e["cjPackedBegin"], e["cjPackedEnd"]):
            cj = hand.cj_packed[g]["cj"]
            if any(8 * e["sci"] <= c < 8 * e["sci"] + 8 for c in cj):
                hand.cj_packed[g]["excl_ind"] = 0
                changed += 1
    assert changed > 0
    real = atoms.atom_indices >= 0
    ic = nb.interaction_const(1.0, 1e-5)
    f64, _, elj64, eel64 = oracle.kernel(hand, atoms, ic, True, "f64")
    for accumulation in (nb.ACCUM_FIXED64, nb.ACCUM_FP64):
        for lds in (0, 1):
            announce(f"diagonal without Newton mask accumulation={accumulation} lds={lds}")
            task = nb.GpuTask(ic, atoms, 1.1, None, accumulation)
            task.set_i_cluster_in_lds(lds)
            task.init_pairlist(hand)
            task.launch()
            f, elj, eel, _ = _fetch_fd(task)
            tol = 1e-10 if accumulation == nb.ACCUM_FP64 else 5e-6
            assert rel_rms(f[real], f64[real]) < tol
            assert elj == pytest.approx(elj64, rel=2e-5)
            assert eel == pytest.approx(eel64, rel=2e-5)


def test_one_hip_runtime_in_process():
    """The library, and the device arrays the tests and the benchmark hand it,
    run on one HIP runtime: the libamdhip64 of /opt/rocm the library is linked
    against (torch's wheel carries another under the same SONAME; nothing
    that drives the library imports torch)."""
    announce("one HIP runtime in the process")
    d = nb.DeviceArray(np.arange(16, dtype=np.float32))
    assert np.array_equal(d.download(), np.arange(16, dtype=np.float32))
    with open("/proc/self/maps") as fh:
        paths = {line.split()[-1] for line in fh if "libamdhip64" in line}
    real = {os.path.realpath(p) for p in paths}
    assert len(real) == 1, real
    assert next(iter(real)).startswith(os.path.realpath("/opt/rocm")), real
