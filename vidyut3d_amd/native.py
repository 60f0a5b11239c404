"""ctypes front end of libvidyut3d_amd.so (the C ABI of include/
vidyut3d_amd.h and include/vidyut3d_amd_host.h).

Device memory is owned by small Python handles (``DeviceArray``) that carry
their Array4 extent; every call goes straight to the HIP library.  There is
no CPU fallback: if the library is missing or no GPU is visible the calls
raise.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libvidyut3d_amd.so")
_lib = None

P = C.POINTER
A4P = P(abi.vd_array4)


class VidyutError(RuntimeError):
    def __init__(self, what, status, msg):
        super().__init__("%s failed: %s (%s)" % (what, abi.STATUS_NAMES.get(status, status), msg))
        self.status = status


def lib():
    """Loads the in-tree HIP library (built by vidyut3d_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libvidyut3d_amd.so is not built: run __graft_entry__.build() or "
                               "python vidyut3d_amd/build.py")
        L = C.CDLL(LIBPATH)
        L.vd_last_error.restype = C.c_char_p
        L.vd_version.restype = C.c_char_p
        L.vd_ctx_stream.restype = C.c_void_p
        L.vd_sim_time.restype = C.c_double
        L.vd_sim_time.argtypes = [C.c_void_p]
        L.vd_sim_dt.restype = C.c_double
        L.vd_sim_dt.argtypes = [C.c_void_p]
        L.vd_sim_mg_iterations.restype = C.c_longlong
        L.vd_sim_mg_iterations.argtypes = [C.c_void_p]
        L.vd_mech_find_id.argtypes = [C.c_void_p, C.c_char_p]
        L.vd_ctx_case_stride.restype = C.c_longlong
        L.vd_ctx_case_stride.argtypes = [C.c_void_p]
        L.vd_ctx_ncase.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def check(st, what, allow=(abi.VD_OK,)):
    if st not in allow:
        raise VidyutError(what, st, lib().vd_last_error().decode(errors="replace"))
    return st


class Device:
    """A vd_ctx (one HIP stream + workspaces) on one GPU.

    Arrays, mechanisms and sims made on a device register with it; close()
    releases the ones still alive before it destroys the context, so that a
    handle outliving its device (held by a traceback, say) frees nothing
    through a destroyed context."""

    def __init__(self, device=0):
        self.ptr = C.c_void_p()
        self._deps = weakref.WeakSet()
        check(lib().vd_ctx_create(int(device), C.byref(self.ptr)), "vd_ctx_create")
        self.ncase = 1

    def adopt(self, obj):
        self._deps.add(obj)
        return obj

    def close(self):
        if self.ptr:
            for o in list(self._deps):
                o.release()
            lib().vd_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().vd_ctx_sync(self.ptr), "vd_ctx_sync")

    def release_workspaces(self):
        check(lib().vd_ctx_release_workspaces(self.ptr), "vd_ctx_release_workspaces")

    def alloc(self, nbytes):
        p = C.c_void_p()
        check(lib().vd_malloc(self.ptr, C.c_size_t(int(nbytes)), C.byref(p)), "vd_malloc")
        return p.value

    def free(self, p):
        lib().vd_free(self.ptr, C.c_void_p(p))

    def upload(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().vd_memcpy_h2d(self.ptr, C.c_void_p(dptr), arr.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(arr.nbytes)), "vd_memcpy_h2d")

    def download(self, dptr, out):
        check(lib().vd_memcpy_d2h(self.ptr, out.ctypes.data_as(C.c_void_p), C.c_void_p(dptr),
                                  C.c_size_t(out.nbytes)), "vd_memcpy_d2h")
        return out

    def set_kernel_timing(self, family):
        check(lib().vd_ctx_set_kernel_timing(self.ptr, int(family)), "vd_ctx_set_kernel_timing")

    def set_kernel_timing_mask(self, families):
        mask = 0
        for f in families:
            mask |= 1 << int(f)
        check(lib().vd_ctx_set_kernel_timing_mask(self.ptr, C.c_uint(mask)), "vd_ctx_set_kernel_timing_mask")

    def kernel_timing_family(self, family):
        """(total ms, launches, canonical bytes) of one timed family"""
        ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
        check(lib().vd_ctx_kernel_timing_family(self.ptr, int(family), C.byref(ms), C.byref(n), C.byref(by)),
              "vd_ctx_kernel_timing_family")
        return ms.value, n.value, by.value

    def kernel_timing(self):
        ms, n, cells = C.c_double(), C.c_longlong(), C.c_double()
        check(lib().vd_ctx_kernel_timing(self.ptr, C.byref(ms), C.byref(n), C.byref(cells)),
              "vd_ctx_kernel_timing")
        return ms.value, n.value, cells.value


class BatchDevice(Device):
    """A batched vd_ctx (vd_ctx_create_batched): ncase same-shape cases,
    each in its own region of case_bytes device memory, stepped together."""

    def __init__(self, device, ncase, case_bytes):
        self.ptr = C.c_void_p()
        self._deps = weakref.WeakSet()
        check(lib().vd_ctx_create_batched(int(device), int(ncase), C.c_size_t(int(case_bytes)), C.byref(self.ptr)),
              "vd_ctx_create_batched")
        self.ncase = int(ncase)

    def set_active(self, cases):
        cases = list(cases)
        check(lib().vd_ctx_set_active(self.ptr, len(cases), (C.c_int * max(len(cases), 1))(*cases)),
              "vd_ctx_set_active")

    def set_case_voltage(self, c, amp_1, amp_2, use=True):
        check(lib().vd_ctx_set_case_voltage(self.ptr, int(c), int(bool(use)), C.c_double(amp_1),
                                            C.c_double(amp_2)), "vd_ctx_set_case_voltage")

    def arena_used(self):
        n = C.c_size_t()
        check(lib().vd_ctx_arena_used(self.ptr, C.byref(n)), "vd_ctx_arena_used")
        return n.value

    def case_solve_info(self, c):
        info = abi.vd_solve_info()
        check(lib().vd_ctx_case_solve_info(self.ptr, int(c), C.byref(info)), "vd_ctx_case_solve_info")
        return info

    def case_time_scales(self, c):
        out = (C.c_double * 3)()
        check(lib().vd_ctx_case_time_scales(self.ptr, int(c), out), "vd_ctx_case_time_scales")
        return list(out)


class DeviceArray:
    """Device Array4 of shape (ncomp, nz, ny, nx) over begin/end."""

    def __init__(self, dev, begin, end, ncomp, host=None, dtype=np.float64):
        self.dev, self.begin, self.end, self.ncomp = dev, list(begin), list(end), ncomp
        self.shape = abi.shape_of(begin, end, ncomp)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = dev.alloc(self.nbytes)
        dev.adopt(self)
        if host is not None:
            self.upload(host)
        else:
            check(lib().vd_memset_zero(dev.ptr, C.c_void_p(self.ptr), C.c_size_t(self.nbytes)),
                  "vd_memset_zero")

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.shape == self.shape, (host.shape, self.shape)
        self.dev.upload(self.ptr, host)

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        return self.dev.download(self.ptr, out)

    def view(self):
        return abi.array4_view(self.ptr, self.begin, self.end, self.ncomp)

    def tag_view(self):
        return abi.tag_view(self.ptr, self.begin, self.end)

    def release(self):
        if self.ptr and self.dev.ptr:
            self.dev.free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Mech:
    def __init__(self, dev, desc):
        self.dev, self.desc = dev, desc
        self.ptr = C.c_void_p()
        check(lib().vd_mech_create(dev.ptr, C.byref(desc), C.byref(self.ptr)), "vd_mech_create")
        dev.adopt(self)

    def find_id(self, name):
        return lib().vd_mech_find_id(self.ptr, name.encode())

    def specialise(self, enable=True):
        """Use (True) the mechanism-specialised source kernel or (False) the
        generic table kernel."""
        check(lib().vd_mech_specialise(self.ptr, 1 if enable else 0), "vd_mech_specialise")

    @property
    def specialised(self):
        return bool(lib().vd_mech_is_specialised(self.ptr))

    def release(self):
        if self.ptr and self.dev.ptr:
            lib().vd_mech_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Grid:
    """Device arrays of one single-box level."""

    def __init__(self, dev, geom):
        self.dev, self.geom = dev, geom
        self.box = abi.domain_box(geom)
        self.dim = geom.dim

    def extent(self, ng=0, face=None):
        return abi.grown_extent(self.box, ng, self.dim, face)

    def array(self, ncomp, ng=0, face=None, host=None):
        b, e = self.extent(ng, face)
        return DeviceArray(self.dev, b, e, ncomp, host)

    def faces(self):
        dims = [0, 1, 2] if self.dim == 3 else [0, 1, 0]
        return [self.array(1, 0, d) for d in dims]

    @staticmethod
    def face_views(fs):
        return (abi.vd_array4 * 3)(*[f.view() for f in fs])


def chop(geom, max_grid_size):
    """The domain cut into boxes of at most max_grid_size cells per
    direction (the last piece of a direction takes the remainder), as
    (lo, hi) pairs in x-fastest order."""
    cuts = []
    for d in range(3):
        if d < geom.dim:
            lo, hi = geom.dom_lo[d], geom.dom_hi[d]
            cuts.append([(a, min(a + max_grid_size - 1, hi)) for a in range(lo, hi + 1, max_grid_size)])
        else:
            cuts.append([(0, 0)])
    return [((x[0], y[0], z[0]), (x[1], y[1], z[1])) for z in cuts[2] for y in cuts[1] for x in cuts[0]]


class Level:
    """A single-level domain held in several boxes (a BoxArray, e.g. level 0
    chopped by amr.max_grid_size), with per-box device arrays."""

    def __init__(self, dev, geom, boxes):
        self.dev, self.geom, self.dim = dev, geom, geom.dim
        self.boxes = [abi.make_box(lo, hi) for lo, hi in boxes]
        self.cboxes = (abi.vd_box * len(self.boxes))(*self.boxes)
        self.whole = Grid(dev, geom)

    def arrays(self, ncomp, ng=0, host=None, host_ng=0):
        """One array per box, grown by ng; with `host` (a domain array grown
        by host_ng >= ng), each box's part of it."""
        out = []
        for b in self.boxes:
            be, en = abi.grown_extent(b, ng, self.dim)
            h = None
            if host is not None:
                sl = tuple(slice(be[d] - self.geom.dom_lo[d] + (host_ng if d < self.dim else 0),
                                 en[d] - self.geom.dom_lo[d] + (host_ng if d < self.dim else 0)) for d in (2, 1, 0))
                h = host[(slice(None),) + sl]
            out.append(DeviceArray(self.dev, be, en, ncomp, h))
        return out

    def box_grid(self, b):
        """A Grid whose box is box b (for the per-box entry points)."""
        g = Grid(self.dev, self.geom)
        g.box = self.boxes[b]
        return g

    @staticmethod
    def views(arrs):
        return (abi.vd_array4 * len(arrs))(*[a.view() for a in arrs])


def fill_patch_boxes(level, src, dst, scomp=0, ncomp=None):
    ncomp = src[0].ncomp if ncomp is None else ncomp
    check(lib().vd_fill_patch_boxes(level.dev.ptr, len(level.boxes), Level.views(src), Level.views(dst),
                                    level.cboxes, C.byref(level.geom), int(scomp), int(ncomp)),
          "vd_fill_patch_boxes")


def level_gather(level, parts, whole, scomp=0, ncomp=None):
    ncomp = parts[0].ncomp if ncomp is None else ncomp
    w = whole.view()
    check(lib().vd_level_gather(level.dev.ptr, len(level.boxes), Level.views(parts), level.cboxes, C.byref(w),
                                C.byref(level.geom), int(scomp), int(ncomp)), "vd_level_gather")


def level_scatter(level, whole, parts, scomp=0, ncomp=None):
    ncomp = parts[0].ncomp if ncomp is None else ncomp
    w = whole.view()
    check(lib().vd_level_scatter(level.dev.ptr, len(level.boxes), C.byref(w), Level.views(parts), level.cboxes,
                                 C.byref(level.geom), int(scomp), int(ncomp)), "vd_level_scatter")


# --------------------------------------------------------- entry points
def fill_patch(grid, src, dst, scomp=0, ncomp=None):
    ncomp = src.ncomp if ncomp is None else ncomp
    s, d = src.view(), dst.view()
    check(lib().vd_fill_patch(grid.dev.ptr, C.byref(s), C.byref(d), C.byref(grid.box),
                              C.byref(grid.geom), int(scomp), int(ncomp)), "vd_fill_patch")


def transport_flux(mech, prm, grid, sb, specid, bclo, bchi, flux, dt=1e-12):
    check(lib().vd_compute_scalar_transport_flux(
        grid.dev.ptr, mech.ptr, C.byref(sb.view()), Grid.face_views(flux), C.byref(grid.box),
        C.byref(grid.geom), abi.ivec(bclo), abi.ivec(bchi), C.byref(prm), C.c_double(0.0),
        C.c_double(dt), int(specid)), "vd_compute_scalar_transport_flux")


def dsdt(grid, flux, rxn, out, specid):
    check(lib().vd_compute_dsdt(grid.dev.ptr, Grid.face_views(flux), C.byref(rxn.view()),
                                C.byref(out.view()), C.byref(grid.box), C.byref(grid.geom),
                                int(specid)), "vd_compute_dsdt")


def transport_rhs(mech, prm, grid, sb, rxn, out, specid, bclo, bchi, dt=1e-12):
    check(lib().vd_transport_rhs(grid.dev.ptr, mech.ptr, C.byref(sb.view()), C.byref(rxn.view()),
                                 C.byref(out.view()), C.byref(grid.box), C.byref(grid.geom),
                                 abi.ivec(bclo), abi.ivec(bchi), C.byref(prm), C.c_double(0.0),
                                 C.c_double(dt), int(specid)), "vd_transport_rhs")


def rxn_source(mech, prm, grid, sb, out):
    check(lib().vd_update_rxnsrc(grid.dev.ptr, mech.ptr, C.byref(sb.view()), C.byref(out.view()),
                                 C.byref(grid.box), C.byref(prm), C.c_double(0.0)),
          "vd_update_rxnsrc")


def elec_energy_source(mech, prm, grid, sb, rxn, efield, gradne, ds, phi_new):
    check(lib().vd_compute_elecenergy_source(
        grid.dev.ptr, mech.ptr, C.byref(sb.view()), C.byref(rxn.view()), Grid.face_views(efield),
        Grid.face_views(gradne), C.byref(ds.view()), C.byref(phi_new.view()), C.byref(grid.box),
        C.byref(grid.geom), C.byref(prm)), "vd_compute_elecenergy_source")


def solve_potential(mech, prm, grid, sb, phi_new, efield, time=0.0):
    info = abi.vd_solve_info()
    check(lib().vd_solve_potential(grid.dev.ptr, mech.ptr, C.byref(sb.view()),
                                   C.byref(phi_new.view()), Grid.face_views(efield),
                                   C.byref(grid.box), C.byref(grid.geom), C.byref(prm),
                                   C.c_double(time), C.byref(info)),
          "vd_solve_potential", (abi.VD_OK, abi.VD_ENOCONV))
    return info


def implicit_solve(mech, prm, grid, spec, sb, phi_old, dsdt_expl, phi_new, grad, bclo, bchi, dt):
    info = abi.vd_solve_info()
    check(lib().vd_implicit_solve_scalar(
        grid.dev.ptr, mech.ptr, int(spec), C.byref(sb.view()), C.byref(phi_old.view()),
        C.byref(dsdt_expl.view()), C.byref(phi_new.view()), Grid.face_views(grad),
        C.byref(grid.box), C.byref(grid.geom), abi.ivec(bclo), abi.ivec(bchi), C.byref(prm),
        C.c_double(0.0), C.c_double(dt), C.byref(info)),
        "vd_implicit_solve_scalar", (abi.VD_OK, abi.VD_ENOCONV))
    return info


def time_scales(mech, prm, grid, phi):
    out = (C.c_double * 3)()
    check(lib().vd_find_time_scales(grid.dev.ptr, mech.ptr, C.byref(phi.view()),
                                    C.byref(grid.box), C.byref(grid.geom), C.byref(prm), out),
          "vd_find_time_scales")
    return list(out)


def error_est(grid, state1, tags, comps, thr, gthr, tagval=1):
    n = len(comps)
    tv = tags.tag_view()
    check(lib().vd_error_est(grid.dev.ptr, C.byref(state1.view()), C.byref(tv), C.byref(grid.box),
                             n, (C.c_int * n)(*comps), (C.c_double * n)(*thr),
                             (C.c_double * n)(*gthr), C.c_char(tagval)), "vd_error_est")


def mlmg_solve(grid, prob, phi, rhs):
    info = abi.vd_solve_info()
    check(lib().vd_mlmg_solve(grid.dev.ptr, C.byref(prob), C.byref(phi.view()),
                              C.byref(rhs.view()), C.byref(grid.box), C.byref(grid.geom),
                              C.byref(info)), "vd_mlmg_solve", (abi.VD_OK, abi.VD_ENOCONV))
    return info


class Sim:
    """vd_sim: the C++ Vidyut driver (device-resident state)."""

    def __init__(self, dev, case):
        self.dev, self.case = dev, case
        self.ptr = C.c_void_p()
        check(lib().vd_sim_create(dev.ptr, C.byref(case.mech), C.byref(case.geom),
                                  C.byref(case.params), C.byref(self.ptr)), "vd_sim_create")
        dev.adopt(self)
        self.shape = case.state.shape
        self.set_state(case.state, 0.0)

    def set_state(self, st, time=0.0):
        st = np.ascontiguousarray(st, dtype=np.float64)
        assert st.shape == self.shape
        check(lib().vd_sim_set_state(self.ptr, st.ctypes.data_as(C.c_void_p), C.c_double(time)),
              "vd_sim_set_state")

    def get_state(self):
        out = np.empty(self.shape)
        check(lib().vd_sim_get_state(self.ptr, out.ctypes.data_as(C.c_void_p)), "vd_sim_get_state")
        return out

    def evolve(self, nsteps, dt=None, ncorr=1):
        dt = self.case.dt if dt is None else dt
        check(lib().vd_sim_evolve(self.ptr, int(nsteps), C.c_double(dt), int(ncorr)),
              "vd_sim_evolve", (abi.VD_OK, abi.VD_ENOCONV))

    # ---- batched contexts: one state per case
    def set_state_case(self, c, st):
        st = np.ascontiguousarray(st, dtype=np.float64)
        assert st.shape == self.shape
        check(lib().vd_sim_set_state_case(self.ptr, int(c), st.ctypes.data_as(C.c_void_p)),
              "vd_sim_set_state_case")

    def get_state_case(self, c):
        out = np.empty(self.shape)
        check(lib().vd_sim_get_state_case(self.ptr, int(c), out.ctypes.data_as(C.c_void_p)),
              "vd_sim_get_state_case")
        return out

    @property
    def time(self):
        return lib().vd_sim_time(self.ptr)

    @property
    def dt(self):
        """the last step's dt"""
        return lib().vd_sim_dt(self.ptr)

    @property
    def mg_iterations(self):
        return lib().vd_sim_mg_iterations(self.ptr)

    def write_plotfile(self, path):
        check(lib().vd_sim_write_plotfile(self.ptr, path.encode()), "vd_sim_write_plotfile")

    def write_checkpoint(self, path):
        """WriteCheckpointFile: Header + Level_0/phi (VisMF) of the state."""
        check(lib().vd_sim_write_checkpoint(self.ptr, path.encode()), "vd_sim_write_checkpoint")

    def read_checkpoint(self, path):
        """ReadCheckpointFile: restores the state, step, dt and time."""
        check(lib().vd_sim_read_checkpoint(self.ptr, path.encode()), "vd_sim_read_checkpoint")

    @property
    def step(self):
        return lib().vd_sim_step(self.ptr)

    # ---- AMR level 1 (the step driver advances level 0 only)
    def regrid(self, comps, thr, gthr, prm):
        """Vidyut::regrid(0, time): tags of level 0 -> level-1 grids
        (MakeNewLevelFromCoarse / RemakeLevel / ClearLevel); returns the
        level-1 boxes."""
        n = len(comps)
        nb = C.c_int()
        check(lib().vd_sim_regrid(self.ptr, n, (C.c_int * n)(*comps), (C.c_double * n)(*thr),
                                  (C.c_double * n)(*gthr), C.byref(prm), C.byref(nb)), "vd_sim_regrid")
        return self.level_boxes(1)

    def level_boxes(self, lev):
        n = C.c_int()
        lib().vd_sim_level_boxes(self.ptr, int(lev), None, 0, C.byref(n))
        out = (abi.vd_box * max(n.value, 1))()
        check(lib().vd_sim_level_boxes(self.ptr, int(lev), out, max(n.value, 1), C.byref(n)),
              "vd_sim_level_boxes")
        return [out[i] for i in range(n.value)]

    def fine_shape(self, box):
        b, e = abi.grown_extent(box, 0, self.case.geom.dim)
        return abi.shape_of(b, e, self.shape[0])

    def get_fine(self, b, box):
        out = np.empty(self.fine_shape(box))
        check(lib().vd_sim_get_fine(self.ptr, int(b), out.ctypes.data_as(C.c_void_p)), "vd_sim_get_fine")
        return out

    def set_fine(self, b, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        check(lib().vd_sim_set_fine(self.ptr, int(b), host.ctypes.data_as(C.c_void_p)), "vd_sim_set_fine")

    def average_down(self):
        check(lib().vd_sim_average_down(self.ptr), "vd_sim_average_down")

    def error_est(self, comps, thr, gthr):
        n = len(comps)
        out = np.zeros(self.shape[1:], dtype=np.int8)
        check(lib().vd_sim_error_est(self.ptr, n, (C.c_int * n)(*comps), (C.c_double * n)(*thr),
                                     (C.c_double * n)(*gthr), out.ctypes.data_as(C.c_char_p)),
              "vd_sim_error_est")
        return out

    def close(self):
        if self.ptr and self.dev.ptr:
            lib().vd_sim_destroy(self.ptr)
        self.ptr = C.c_void_p()

    release = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------ level transfers (AMR)
def _ratio(ratio):
    return (C.c_int * 3)(*[int(x) for x in ratio])


def average_down(dev, fine, fbox, crse, scomp, ncomp, ratio, dim):
    """Vidyut::AverageDown (Source/GridWork.cpp:94-108) for one fine box:
    the coarse cells under fbox become the mean of their children."""
    f, c = fine.view(), crse.view()
    check(lib().vd_average_down(dev.ptr, C.byref(f), C.byref(fbox), C.byref(c), int(scomp), int(ncomp),
                                _ratio(ratio), int(dim)), "vd_average_down")


def average_down_faces(dev, fine_faces, fbox, crse_faces, ratio, dim):
    """average_down_faces of the species fluxes (Source/ScalarSolve.cpp:88-94)."""
    fv = (abi.vd_array4 * 3)(*[a.view() for a in fine_faces])
    cv = (abi.vd_array4 * 3)(*[a.view() for a in crse_faces])
    check(lib().vd_average_down_faces(dev.ptr, fv, C.byref(fbox), cv, _ratio(ratio), int(dim)),
          "vd_average_down_faces")


def fill_patch_two_levels(dev, crse, cgeom, fines, fboxes, dst, fgeom, ratio, scomp=0, ncomp=None):
    """Vidyut::FillPatch at lev > 0 (FillPatchTwoLevels, cell_cons_interp);
    with no fine boxes, Vidyut::FillCoarsePatch (Source/GridWork.cpp:123-164)."""
    ncomp = dst.ncomp - scomp if ncomp is None else ncomp
    nf = len(fboxes)
    fv = (abi.vd_array4 * max(nf, 1))(*[a.view() for a in fines])
    bv = (abi.vd_box * max(nf, 1))(*fboxes)
    cv, dv = crse.view(), dst.view()
    check(lib().vd_fill_patch_two_levels(dev.ptr, C.byref(cv), C.byref(cgeom), fv, bv, nf, C.byref(dv),
                                         C.byref(fgeom), _ratio(ratio), int(scomp), int(ncomp)),
          "vd_fill_patch_two_levels")


# ------------------------------------------------------------- regrid
def regrid_params(ref_ratio=2, blocking_factor=2, max_grid_size=32, n_error_buf=1, grid_eff=0.7):
    """amr.* keys of the new level (scalars apply to every direction)."""
    p = abi.vd_regrid_params()
    for name, v in (("ref_ratio", ref_ratio), ("blocking_factor", blocking_factor),
                    ("max_grid_size", max_grid_size)):
        v = [v] * 3 if np.isscalar(v) else list(v) + [1] * (3 - len(v))
        getattr(p, name)[:] = [int(x) for x in v]
    p.n_error_buf = int(n_error_buf)
    p.grid_eff = float(grid_eff)
    return p


def make_new_grids(tags, geom, prm):
    """AmrMesh::MakeNewGrids on the tags of a single-box level (host code):
    tags is an int8 array (nz, ny, nx) over the domain; returns the new
    level's boxes as a list of vd_box (fine index space)."""
    tags = np.ascontiguousarray(tags, dtype=np.int8)
    dom = abi.domain_box(geom)
    per = (C.c_int * 3)(*[int(geom.is_periodic[d]) for d in range(3)])
    n = C.c_int()
    cap = 64
    while True:
        out = (abi.vd_box * cap)()
        st = lib().vd_make_new_grids(tags.ctypes.data_as(C.c_char_p), C.byref(dom), per, int(geom.dim),
                                     C.byref(prm), out, cap, C.byref(n))
        if st == abi.VD_OK:
            return [out[i] for i in range(n.value)]
        if n.value > cap:
            cap = n.value
            continue
        check(st, "vd_make_new_grids")
