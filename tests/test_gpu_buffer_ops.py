"""GPU buffer ops around the kernel: coordinates from a device array in atom
order into xq, and the force reduction into a device array in atom order.
Device arrays come from the library's own HIP runtime (memory only); the ops
run through the C ABI."""
import numpy as np
import pytest

import nbnxm_amd as nb
from conftest import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    s = nb.systems.config_water_small()
    ic = nb.interaction_const(1.0, 1e-5)
    search, plist, atoms = nb.build(s, 1.1, nb.VDW_CUT_COMB_GEOM)
    # Coordinates as the kernel sees them: the wrapped FP32 values of the grid.
    xw = np.zeros((s.n, 3), dtype=np.float32)
    xq = atoms.xq.reshape(-1, 4)
    real = atoms.atom_indices >= 0
    xw[atoms.atom_indices[real]] = xq[real, :3]
    return dict(s=s, ic=ic, plist=plist, atoms=atoms, xw=xw, real=real)


def _task(st, accumulation=nb.ACCUM_FIXED64):
    task = nb.GpuTask(st["ic"], st["atoms"], 1.1, None, accumulation)
    task.init_pairlist(st["plist"])
    task.init_x_to_nbat_x(st["s"].n, st["atoms"].atom_indices)
    return task


@pytest.mark.parametrize("accumulation", [nb.ACCUM_FIXED64, nb.ACCUM_FLOAT])
def test_x_to_nbat_x_and_force_reduction(setup, accumulation):
    announce(f"buffer ops water9k accumulation={accumulation}")
    st = setup
    n = st["s"].n
    # Reference: host xq, cpyback, host scatter to atom order.
    ref_task = _task(st, accumulation)
    ref_task.launch(compute_energy=True, compute_virial=True)
    f_ref, _, elj_ref, eel_ref, fs_ref = ref_task.fetch(want_double=False)
    vir_ref = ref_task.virial()
    f_ref = st["atoms"].to_original_order(f_ref, n).astype(np.float32)

    # Device path: perturb xq first so that x_to_nbat_x must restore it.
    task = _task(st, accumulation)
    task.copy_xq(np.where(np.arange(st["atoms"].xq.size) % 4 == 3, st["atoms"].xq, 0.0).astype(np.float32))
    d_x = nb.DeviceArray(st["xw"].reshape(-1))
    d_f = nb.DeviceArray.zeros(3 * n)
    task.clear_outputs()
    task.x_to_nbat_x(d_x.ptr)
    task.launch(compute_energy=True, compute_virial=True)
    task.add_nbat_f_to_f(d_f.ptr)
    # Energies and virial still come through cpyback/wait without forces.
    import ctypes
    lib = nb.load()
    nb._lib.check(lib.nbnxm_gpu_launch_cpyback(task._h, None, None), "cpyback")
    e_lj, e_el = ctypes.c_double(0.0), ctypes.c_double(0.0)
    nb._lib.check(lib.nbnxm_gpu_wait_finish_task(task._h, ctypes.byref(e_lj), ctypes.byref(e_el), None), "wait")
    f = d_f.download().reshape(-1, 3)
    if accumulation == nb.ACCUM_FIXED64:
        assert np.array_equal(f, f_ref)
        assert e_lj.value == elj_ref and e_el.value == eel_ref
        assert np.array_equal(task.virial(), vir_ref)
    else:
        # Float atomics: order-dependent in the last bits.
        assert np.abs(f - f_ref).max() <= 1e-5 * np.abs(f_ref).max()
        assert np.abs(task.virial() - vir_ref).max() <= 1e-5 * np.abs(vir_ref).max()


def test_force_reduction_accumulate_and_add(setup):
    announce("buffer ops accumulate + rvec add")
    st = setup
    n = st["s"].n
    ref_task = _task(st)
    ref_task.launch(compute_energy=False, compute_virial=False)
    f_ref, _, _, _, _ = ref_task.fetch(want_double=False)
    f_ref = st["atoms"].to_original_order(f_ref, n).astype(np.float32)
    rng = np.random.default_rng(3)
    f0 = rng.normal(0, 100, (n, 3)).astype(np.float32)
    fadd = rng.normal(0, 100, (n, 3)).astype(np.float32)
    task = _task(st)
    d_x = nb.DeviceArray(st["xw"].reshape(-1))
    d_f = nb.DeviceArray(f0.reshape(-1))
    d_add = nb.DeviceArray(fadd.reshape(-1))
    task.x_to_nbat_x(d_x.ptr)
    task.launch(compute_energy=False, compute_virial=False)
    task.add_nbat_f_to_f(d_f.ptr, d_add.ptr, accumulate=True)
    import ctypes
    lib = nb.load()
    nb._lib.check(lib.nbnxm_gpu_wait_finish_task(task._h, None, None, None), "wait")
    f = d_f.download().reshape(-1, 3)
    expect = (f0 + f_ref) + fadd  # the reference kernel's order: (f + nbnxm) + rvec
    assert np.array_equal(f, expect)
    # The forces of a launch are taken once.
    with pytest.raises(nb.NbnxmError, match="pending"):
        task.add_nbat_f_to_f(d_f.ptr)
    with pytest.raises(nb.NbnxmError, match="taken"):
        task.fetch(want_double=False)
