"""Host-side mirror of the reference's nbnxm objects on top of the C ABI.

  PairSearch    -- nonbonded_verlet_t::putAtomsOnGrid + constructPairlist for
                   the GPU list type (src/gromacs/nbnxm/nbnxm.h:266-515)
  build_atomdata-- nbnxn_atomdata_t kernel inputs in grid order
                   (src/gromacs/nbnxm/atomdata.cpp:261-586,629-790)
  GpuTask       -- the Nbnxm::gpu_* task object (src/gromacs/nbnxm/nbnxm_gpu.h)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ACCUM_FIXED64, CJ_DTYPE, EXCL_DTYPE, NUM_SHIFTS, SCI_DTYPE, VDW_CUT_COMB_GEOM,
                   VDW_CUT_COMB_LB, VDW_CUT_TYPE_TABLE, Atomdata, InteractionConst, PairlistGpu,
                   PairlistParams, check, ptr)


def interaction_const(rc=1.0, ewald_rtol=1e-5, epsilon_r=1.0):
    """Ewald (analytical) + LJ cut-off with PotShift on both."""
    ic = InteractionConst()
    check(_lib.load().nbnxm_init_interaction_const(rc, ewald_rtol, epsilon_r, ctypes.byref(ic)), "init_interaction_const")
    return ic


def interaction_const_ex(rcoulomb, coulomb="ewald", vdw_modifier="potshift", lj_pme=False, rvdw=0.0,
                         rvdw_switch=0.0, ewald_rtol=1e-5, ewald_rtol_lj=1e-3, epsilon_r=1.0, epsilon_rf=0.0,
                         plain_cutoff=False, tabulated=False):
    """Any kernel flavour: coulomb "ewald" | "rf"; vdw_modifier "potshift" |
    "forceswitch" | "potswitch"; lj_pme adds the LJ-PME grid correction with
    the geometric (True or "geom") or Lorentz-Berthelot ("lb") grid rule;
    rvdw < rcoulomb gives the twin cut-off; tabulated selects the tabulated
    Ewald correction."""
    o = _lib.IcOptions()
    o.eeltype = {"ewald": _lib.ELEC_EWALD, "rf": _lib.ELEC_RF}[coulomb]
    o.vdwModifier = {"potshift": _lib.VDW_MOD_POTSHIFT, "forceswitch": _lib.VDW_MOD_FORCESWITCH,
                     "potswitch": _lib.VDW_MOD_POTSWITCH}[vdw_modifier]
    o.ljPme = {False: _lib.LJPME_NONE, True: _lib.LJPME_GEOM, "geom": _lib.LJPME_GEOM,
               "lb": _lib.LJPME_LB}[lj_pme]
    o.rcoulomb = rcoulomb
    o.rvdw = rvdw
    o.rvdw_switch = rvdw_switch
    o.ewaldRtol = ewald_rtol
    o.ewaldRtolLj = ewald_rtol_lj
    o.epsilonR = epsilon_r
    o.epsilonRf = epsilon_rf
    o.plainCutoff = int(bool(plain_cutoff))
    o.tabulatedEwald = int(bool(tabulated))
    ic = InteractionConst()
    check(_lib.load().nbnxm_init_interaction_const_ex(ctypes.byref(o), ctypes.byref(ic)), "init_interaction_const_ex")
    return ic


def is_default_flavour(ic):
    """Ewald + LJ cut-off with potential shift and rvdw == rcoulomb."""
    return (ic.eeltype == _lib.ELEC_EWALD and not ic.tabulatedEwald and ic.vdwModifier == _lib.VDW_MOD_POTSHIFT
            and not ic.ljPme and ic.rvdw == ic.rcoulomb)


def ewald_correction_table(beta, rc):
    """(scale, F table FP32, F table FP64) of the tabulated Ewald correction."""
    lib = _lib.load()
    scale = ctypes.c_double(0.0)
    n = lib.nbnxm_ewald_correction_table(beta, rc, None, None, 0, ctypes.byref(scale))
    if n < 0:
        check(n, "ewald_correction_table")
    t = np.zeros(n, dtype=np.float32)
    td = np.zeros(n, dtype=np.float64)
    rc2 = lib.nbnxm_ewald_correction_table(beta, rc, ptr(t), ptr(td), n, ctypes.byref(scale))
    if rc2 != n:
        check(rc2 if rc2 < 0 else -1, "ewald_correction_table")
    return scale.value, t, td


def ewald_coeff(rc, rtol):
    return _lib.load().nbnxm_calc_ewaldcoeff_q(rc, rtol)


def ewald_coeff_lj(rc, rtol):
    return _lib.load().nbnxm_calc_ewaldcoeff_lj(rc, rtol)


class Pairlist:
    """Owned numpy copy of a GPU pair list (NbnxnPairlistGpu, split 1)."""

    def __init__(self, sci, cj_packed, excl, rlist):
        self.sci = np.ascontiguousarray(sci, dtype=SCI_DTYPE)
        self.cj_packed = np.ascontiguousarray(cj_packed, dtype=CJ_DTYPE)
        self.excl = np.ascontiguousarray(excl, dtype=EXCL_DTYPE)
        self.rlist = float(rlist)

    def view(self):
        v = PairlistGpu()
        v.na_ci = 8
        v.rlist = self.rlist
        v.nsci = len(self.sci)
        v.sci = ptr(self.sci) if len(self.sci) else None
        v.ncjPacked = len(self.cj_packed)
        v.cjPacked = ptr(self.cj_packed) if len(self.cj_packed) else None
        v.nexcl = len(self.excl)
        v.excl = ptr(self.excl)
        return v

    @property
    def imask(self):
        return self.cj_packed["imask"]

    def num_cluster_pairs(self):
        return int(sum(bin(int(m)).count("1") for m in self.imask))

    def copy(self):
        return Pairlist(self.sci.copy(), self.cj_packed.copy(), self.excl.copy(), self.rlist)


class PairSearch:
    """Grid + GPU pair list producer."""

    def __init__(self):
        self._lib = _lib.load()
        self._h = self._lib.nbnxm_search_create()
        if not self._h:
            raise MemoryError("nbnxm_search_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_search_free(self._h)
            self._h = None

    def put_atoms_on_grid(self, box, x):
        box = np.ascontiguousarray(box, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.n = x.shape[0]
        check(self._lib.nbnxm_search_put_atoms_on_grid(self._h, ptr(box), self.n, ptr(x)), "put_atoms_on_grid")

    def construct_pairlist(self, rlist, excl_ranges=None, excl_elements=None, max_pairs_per_entry=0):
        r = None if excl_ranges is None else np.ascontiguousarray(excl_ranges, dtype=np.int32)
        e = None if excl_elements is None else np.ascontiguousarray(excl_elements, dtype=np.int32)
        if e is not None and e.size == 0:
            e = np.zeros(1, dtype=np.int32)
        check(self._lib.nbnxm_search_construct_pairlist(self._h, rlist, ptr(r), ptr(e), int(max_pairs_per_entry)), "construct_pairlist")

    def count_pairs(self, rc):
        """Non-excluded pairs (all images) within rc in the current list, FP64."""
        c = ctypes.c_int64(0)
        check(self._lib.nbnxm_search_count_pairs(self._h, float(rc), ctypes.byref(c)), "search_count_pairs")
        return c.value

    def pairlist(self):
        v = PairlistGpu()
        check(self._lib.nbnxm_search_get_pairlist(self._h, ctypes.byref(v)), "get_pairlist")

        def grab(addr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(addr)
            return np.frombuffer(buf, dtype=dtype).copy()

        return Pairlist(grab(v.sci, v.nsci, SCI_DTYPE), grab(v.cjPacked, v.ncjPacked, CJ_DTYPE),
                        grab(v.excl, v.nexcl, EXCL_DTYPE), v.rlist)

    @property
    def num_grid_atoms(self):
        return self._lib.nbnxm_search_num_grid_atoms(self._h)

    def atom_indices(self):
        out = np.empty(self.num_grid_atoms, dtype=np.int32)
        check(self._lib.nbnxm_search_atom_indices(self._h, ptr(out)), "atom_indices")
        return out

    def shift_vec(self):
        out = np.empty(3 * NUM_SHIFTS, dtype=np.float32)
        check(self._lib.nbnxm_search_shift_vec(self._h, ptr(out)), "shift_vec")
        return out

    def fill_xq(self, x=None, q=None):
        out = np.empty(4 * self.num_grid_atoms, dtype=np.float32)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        qa = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        check(self._lib.nbnxm_search_fill_xq(self._h, ptr(xa), ptr(qa), ptr(out)), "fill_xq")
        return out

    def fill_types(self, types, dummy_type):
        out = np.empty(self.num_grid_atoms, dtype=np.int32)
        t = np.ascontiguousarray(types, dtype=np.int32)
        check(self._lib.nbnxm_search_fill_types(self._h, ptr(t), int(dummy_type), ptr(out)), "fill_types")
        return out

    def fill_ljcomb(self, types, per_type_comb):
        out = np.empty(2 * self.num_grid_atoms, dtype=np.float32)
        t = np.ascontiguousarray(types, dtype=np.int32)
        c = np.ascontiguousarray(per_type_comb, dtype=np.float32)
        check(self._lib.nbnxm_search_fill_ljcomb(self._h, ptr(t), ptr(c), ptr(out)), "fill_ljcomb")
        return out


def nbfp_table(c6c12):
    """(ntypes, ntypes, 2) C6/C12 -> flat (6*C6, 12*C12) table with a trailing
    dummy type, as nbnxn_atomdata_t::Params::nbfp (atomdata.cpp:417, atomdata.h:190)."""
    nt = c6c12.shape[0]
    tab = np.zeros((nt + 1, nt + 1, 2), dtype=np.float64)
    tab[:nt, :nt, 0] = 6.0 * c6c12[:, :, 0]
    tab[:nt, :nt, 1] = 12.0 * c6c12[:, :, 1]
    return tab


def comb_params(c6c12, rule):
    """Per-type combination parameters from the nbfp diagonal (atomdata.cpp:296-330):
    geometric sqrt(6C6), sqrt(12C12); Lorentz-Berthelot 0.5*(12C12/6C6)^(1/6), sqrt(6C6^2/12C12)."""
    tab = nbfp_table(c6c12)
    nt = tab.shape[0]
    out = np.zeros((nt, 2), dtype=np.float64)
    for i in range(nt):
        c6, c12 = tab[i, i]
        if rule == "geom":
            out[i] = (np.sqrt(c6), np.sqrt(c12))
        elif c6 > 0 and c12 > 0:
            out[i] = (0.5 * (c12 / c6) ** (1.0 / 6.0), np.sqrt(c6 * c6 / c12))
    return out


class AtomData:
    """Grid-order kernel inputs (owns the arrays referenced by the C view)."""

    def __init__(self, search, system, vdw_type, x=None):
        nt = system.num_types
        self.vdw_type = vdw_type
        self.num_atoms = search.num_grid_atoms
        self.xq = search.fill_xq(x, system.q)
        self.num_types = nt + 1
        self.types = search.fill_types(system.types, nt)
        self.nbfp = nbfp_table(system.c6c12).astype(np.float32).reshape(-1)
        if vdw_type == VDW_CUT_COMB_GEOM:
            comb = comb_params(system.c6c12, "geom")
        elif vdw_type == VDW_CUT_COMB_LB:
            comb = comb_params(system.c6c12, "lb")
        else:
            comb = np.zeros((nt + 1, 2))
        self.lj_comb = search.fill_ljcomb(system.types, comb[:nt].astype(np.float32).reshape(-1))
        self.shift_vec = search.shift_vec()
        self.atom_indices = search.atom_indices()

    def view(self):
        v = Atomdata()
        v.numAtoms = self.num_atoms
        v.xq = ptr(self.xq)
        v.numTypes = self.num_types
        v.type = ptr(self.types)
        v.nbfp = ptr(self.nbfp)
        v.vdwType = self.vdw_type
        v.ljComb = ptr(self.lj_comb)
        v.shiftVec = ptr(self.shift_vec)
        return v

    def to_original_order(self, f_grid, n):
        """Grid-order (3 per slot) -> original atom order (n, 3), as
        atomdata_add_nbat_f_to_f (atomdata.cpp:1224-1262)."""
        f_grid = np.asarray(f_grid).reshape(-1, 3)
        out = np.zeros((n, 3), dtype=f_grid.dtype)
        real = self.atom_indices >= 0
        out[self.atom_indices[real]] = f_grid[real]
        return out


def build(system, rlist, vdw_type=VDW_CUT_TYPE_TABLE, max_pairs_per_entry=0):
    """Grid, list and atom data for one system."""
    search = PairSearch()
    search.put_atoms_on_grid(system.box, system.x)
    search.construct_pairlist(rlist, system.excl_ranges, system.excl_elements, max_pairs_per_entry)
    plist = search.pairlist()
    atoms = AtomData(search, system, vdw_type)
    return search, plist, atoms


class GpuTask:
    """One system on the GPU, driven through the gpu_* entry points."""

    def __init__(self, ic, atoms, rlist_outer, rlist_inner=None, accumulation=ACCUM_FIXED64,
                 dynamic_pruning=True, rolling_parts=1):
        self._lib = _lib.load()
        self.atoms = atoms
        lp = PairlistParams()
        lp.rlistOuter = rlist_outer
        lp.rlistInner = rlist_outer if rlist_inner is None else rlist_inner
        lp.useDynamicPruning = 1 if dynamic_pruning else 0
        lp.numRollingPruningParts = rolling_parts
        h = ctypes.c_void_p()
        check(self._lib.nbnxm_gpu_init(ctypes.byref(ic), ctypes.byref(lp), ctypes.byref(atoms.view()), accumulation, ctypes.byref(h)), "gpu_init")
        self._h = h
        self.accumulation = accumulation
        self.plist = None

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_gpu_free(self._h)
            self._h = None

    def init_pairlist(self, plist):
        self.plist = plist
        check(self._lib.nbnxm_gpu_init_pairlist(self._h, ctypes.byref(plist.view())), "gpu_init_pairlist")

    def copy_xq(self, xq):
        xq = np.ascontiguousarray(xq, dtype=np.float32)
        check(self._lib.nbnxm_gpu_copy_xq_to_gpu(self._h, ptr(xq), xq.size // 4), "gpu_copy_xq_to_gpu")

    def clear_outputs(self, compute_virial=False):
        check(self._lib.nbnxm_gpu_clear_outputs(self._h, int(compute_virial)), "gpu_clear_outputs")

    def launch(self, compute_energy=True, compute_virial=True):
        check(self._lib.nbnxm_gpu_launch_kernel(self._h, int(compute_energy), int(compute_virial)), "gpu_launch_kernel")

    def prune(self, num_parts=1):
        check(self._lib.nbnxm_gpu_launch_kernel_pruneonly(self._h, int(num_parts)), "gpu_launch_kernel_pruneonly")

    def set_waves_per_block(self, w):
        check(self._lib.nbnxm_gpu_set_waves_per_block(self._h, int(w)), "gpu_set_waves_per_block")

    def set_i_cluster_in_lds(self, mode):
        """1: i-cluster data in LDS (4 waves per SIMD), 0: in registers, -1: by entry count."""
        check(self._lib.nbnxm_gpu_set_i_cluster_in_lds(self._h, int(mode)), "gpu_set_i_cluster_in_lds")

    def fetch(self, want_double=True):
        """cpyback + wait: returns (f_float, f_double or None, e_lj, e_el, fshift)."""
        n3 = 3 * self.atoms.num_atoms
        f = np.zeros(n3, dtype=np.float32)
        fd = np.zeros(n3, dtype=np.float64) if (want_double and self.accumulation != _lib.ACCUM_FLOAT) else None
        check(self._lib.nbnxm_gpu_launch_cpyback(self._h, ptr(f), ptr(fd)), "gpu_launch_cpyback")
        elj = ctypes.c_double(0.0)
        eel = ctypes.c_double(0.0)
        fshift = np.zeros(3 * NUM_SHIFTS, dtype=np.float64)
        check(self._lib.nbnxm_gpu_wait_finish_task(self._h, ctypes.byref(elj), ctypes.byref(eel), ptr(fshift)), "gpu_wait_finish_task")
        return f, fd, elj.value, eel.value, fshift

    def min_ci_balanced(self):
        """gpu_min_ci_balanced: the entry count a list producer should split toward."""
        n = self._lib.nbnxm_gpu_min_ci_balanced(self._h)
        if n < 0:
            check(n, "gpu_min_ci_balanced")
        return int(n)

    # GPU buffer ops: device pointers (ints) to float32 arrays in atom order.
    def init_x_to_nbat_x(self, num_atoms, atom_indices):
        ai = np.ascontiguousarray(atom_indices, dtype=np.int32)
        check(self._lib.nbnxm_gpu_init_x_to_nbat_x(self._h, int(num_atoms), ai.size, ptr(ai)), "gpu_init_x_to_nbat_x")

    def x_to_nbat_x(self, d_x):
        check(self._lib.nbnxm_gpu_x_to_nbat_x(self._h, ctypes.c_void_p(d_x)), "gpu_x_to_nbat_x")

    def add_nbat_f_to_f(self, d_f, d_f_to_add=None, accumulate=False):
        check(self._lib.nbnxm_gpu_add_nbat_f_to_f(self._h, ctypes.c_void_p(d_f),
                                                  None if d_f_to_add is None else ctypes.c_void_p(d_f_to_add),
                                                  int(accumulate)), "gpu_add_nbat_f_to_f")

    def try_finish(self):
        """gpu_try_finish_task in Check mode: None while the task runs, else
        (e_lj, e_el, fshift) added as wait_finish_task adds them."""
        elj = ctypes.c_double(0.0)
        eel = ctypes.c_double(0.0)
        fshift = np.zeros(3 * NUM_SHIFTS, dtype=np.float64)
        rc = self._lib.nbnxm_gpu_try_finish_task(self._h, ctypes.byref(elj), ctypes.byref(eel), ptr(fshift))
        if rc < 0:
            check(rc, "gpu_try_finish_task")
        return None if rc == 0 else (elj.value, eel.value, fshift)

    def wait(self):
        """gpu_wait_finish_task: (e_lj, e_el, fshift) of the last launch."""
        elj = ctypes.c_double(0.0)
        eel = ctypes.c_double(0.0)
        fshift = np.zeros(3 * NUM_SHIFTS, dtype=np.float64)
        check(self._lib.nbnxm_gpu_wait_finish_task(self._h, ctypes.byref(elj), ctypes.byref(eel), ptr(fshift)), "gpu_wait_finish_task")
        return elj.value, eel.value, fshift

    def pme_loadbal_update_param(self, ic, rlist_outer=None, rlist_inner=None, rolling_parts=1, dynamic_pruning=True):
        """gpu_pme_loadbal_update_param: new interaction constants (and list radii)."""
        lp = None
        if rlist_outer is not None:
            lp = PairlistParams()
            lp.rlistOuter = rlist_outer
            lp.rlistInner = rlist_outer if rlist_inner is None else rlist_inner
            lp.useDynamicPruning = 1 if dynamic_pruning else 0
            lp.numRollingPruningParts = rolling_parts
        check(self._lib.nbnxm_gpu_pme_loadbal_update_param(self._h, ctypes.byref(ic), None if lp is None else ctypes.byref(lp)),
              "gpu_pme_loadbal_update_param")

    def set_timing(self, enable=True):
        check(self._lib.nbnxm_gpu_set_timing(self._h, int(enable)), "gpu_set_timing")

    def timings(self):
        """gpu_get_timings: the GpuTimings struct, or None when timing is off."""
        t = _lib.GpuTimings()
        rc = self._lib.nbnxm_gpu_get_timings(self._h, ctypes.byref(t))
        if rc < 0:
            check(rc, "gpu_get_timings")
        return t if rc == 1 else None

    def reset_timings(self):
        check(self._lib.nbnxm_gpu_reset_timings(self._h), "gpu_reset_timings")

    def get_f(self):
        """gpu_get_f: device address (int) of the nbat float force buffer."""
        p = ctypes.c_void_p()
        check(self._lib.nbnxm_gpu_get_f(self._h, ctypes.byref(p)), "gpu_get_f")
        return p.value

    def stream(self):
        """The task's hipStream_t as an int."""
        p = ctypes.c_void_p()
        check(self._lib.nbnxm_gpu_get_stream(self._h, ctypes.byref(p)), "gpu_get_stream")
        return p.value

    def virial(self):
        """Single-sum virial (3, 3) of the last launch made with compute_virial."""
        v = np.zeros(9, dtype=np.float64)
        check(self._lib.nbnxm_gpu_get_virial(self._h, ptr(v)), "gpu_get_virial")
        return v.reshape(3, 3)

    def masks(self):
        n = len(self.plist.cj_packed)
        inner = np.zeros(n, dtype=np.uint32)
        outer = np.zeros(n, dtype=np.uint32)
        check(self._lib.nbnxm_gpu_get_masks(self._h, ptr(inner), ptr(outer)), "gpu_get_masks")
        return inner, outer

    def sci_order(self):
        n = len(self.plist.sci)
        order = np.zeros(n, dtype=np.int32)
        count = np.zeros(n, dtype=np.int32)
        check(self._lib.nbnxm_gpu_get_sci_order(self._h, ptr(order), ptr(count)), "gpu_get_sci_order")
        return order, count


class GpuPairSearch:
    """Grids and GPU pair lists of one or many systems built on the device
    (nbnxm_gpu_search_*); each system's list equals PairSearch's."""

    def __init__(self):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.nbnxm_gpu_search_create(ctypes.byref(h)), "gpu_search_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.nbnxm_gpu_search_free(self._h)
            self._h = None

    def build(self, systems, rlist, max_pairs_per_entry=0):
        from .ensemble import system_descs
        descs, keep = system_descs(systems)
        check(self._lib.nbnxm_gpu_search_build(self._h, len(systems), ctypes.cast(descs, ctypes.c_void_p), float(rlist),
                                               int(max_pairs_per_entry)), "gpu_search_build")
        self.rlist = float(rlist)
        del keep

    def num_entries(self):
        c = np.zeros(4, dtype=np.int64)
        check(self._lib.nbnxm_gpu_search_num_entries(self._h, ptr(c)), "gpu_search_num_entries")
        return dict(zip(["grid_slots", "entries", "groups", "excl"], (int(v) for v in c)))

    def system_list(self, k):
        """(Pairlist, slot -> atom map) of system k in single-system numbering."""
        sizes = np.zeros(4, dtype=np.int32)
        check(self._lib.nbnxm_gpu_search_system_list(self._h, int(k), None, None, None, None, ptr(sizes)), "gpu_search_system_list")
        sci = np.zeros(max(int(sizes[1]), 1), dtype=SCI_DTYPE)
        cj = np.zeros(max(int(sizes[2]), 1), dtype=CJ_DTYPE)
        ex = np.zeros(int(sizes[3]), dtype=EXCL_DTYPE)
        ai = np.zeros(int(sizes[0]), dtype=np.int32)
        check(self._lib.nbnxm_gpu_search_system_list(self._h, int(k), ptr(sci), ptr(cj), ptr(ex), ptr(ai), ptr(sizes)), "gpu_search_system_list")
        return Pairlist(sci[:sizes[1]], cj[:sizes[2]], ex, self.rlist), ai
