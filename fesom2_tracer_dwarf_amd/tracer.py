"""Host-side interface to the GPU tracer advection, mirroring the reference's Python wrapper.

``FESOMTracerAdvection`` keeps the method names of python/demo_tracer.py:40-208
(``initialize_mpi``, ``advect_step``, ``get_tracer_values``, ``get_statistics``, ``finalize``,
``finalize_mpi``) and adds array-based initialisation from a :class:`~.configs.Case`, the
do_oce_adv_tra drop-in call, and PyTorch device-tensor exchange.  Every call goes through the C ABI
of libfesomtra.so; there is no CPU path.
"""

from __future__ import annotations

import ctypes
from typing import Dict, Optional

import numpy as np

from . import abi


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class FESOMTracerAdvection:
    """Process-wide tracer advection instance (the library holds one, as the reference does)."""

    def __init__(self):
        self.lib = abi.load()
        self.nl = self.nod2D = self.elem2D = self.edge2D = 0
        self.num_tracers = 0

    # --- reference wrapper API (python/demo_tracer.py) ------------------------------------------
    def initialize_mpi(self, mpi_comm: int = 0) -> int:
        return self.lib.tracer_init_mpi(mpi_comm)

    def initialize(self, restart_path: str = "./fesom_bin_restart") -> int:
        return self.lib.tracer_init(restart_path.encode())

    def advect_step(self, dt: float = 1.0e-3) -> int:
        return self.lib.tracer_advect_step(dt)

    def get_tracer_values(self, tracer_id: int = 1) -> np.ndarray:
        nz, nn = ctypes.c_int(), ctypes.c_int()
        abi.check(self.lib.tracer_get_size(ctypes.byref(nz), ctypes.byref(nn)), "tracer_get_size")
        out = np.zeros((nn.value, nz.value), dtype=np.float64)  # == Fortran (nz, nn)
        abi.check(self.lib.tracer_get_values_by_id(tracer_id, abi.dptr(out), nz.value, nn.value),
                  "tracer_get_values")
        return out

    def get_statistics(self) -> Dict[str, float]:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        abi.check(self.lib.tracer_get_stats(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                  "tracer_get_stats")
        return {"min": a.value, "max": b.value, "sum": c.value}

    def finalize(self, output_path: str = "./fesom_bin_restart") -> int:
        return self.lib.tracer_finalize(output_path.encode())

    def finalize_mpi(self) -> int:
        return self.lib.tracer_finalize_mpi()

    # --- array-based initialisation ---------------------------------------------------------------
    def setup_case(self, case, upload: bool = True, mesh_files: Optional[str] = None) -> None:
        """Hand a generated case (mesh, velocities, tracers, schemes) to the library.  With
        ``mesh_files`` the library reads the mesh itself from that FESOM2 mesh directory
        (tracer_load_mesh_from_files) instead of taking ``case.mesh``'s arrays; ``case.mesh`` must
        be the same mesh (its sizes and element order index the velocities and tracers)."""
        L = self.lib
        m = case.mesh
        self.nl, self.nod2D, self.elem2D, self.edge2D = m.nl, m.nod2D, m.elem2D, m.edge2D
        self.num_tracers = case.ntracers
        abi.check(L.tracer_init_mpi(0), "tracer_init_mpi")
        self._keep = []

        def I(a):
            a = _i32(a)
            self._keep.append(a)
            return abi.iptr(a)

        def D(a):
            a = _f64(a)
            self._keep.append(a)
            return abi.dptr(a)

        if mesh_files is not None:
            abi.check(L.tracer_load_mesh_from_files(str(mesh_files).encode()), "tracer_load_mesh_from_files")
        else:
            abi.check(L.tracer_set_mesh_dims(m.nl, m.nod2D, m.edge2D, m.elem2D), "tracer_set_mesh_dims")
            abi.check(L.tracer_set_partition(m.nod2D, 0), "tracer_set_partition")
            abi.check(L.tracer_set_coordinates(D(m.coord_nod2D), m.nod2D), "tracer_set_coordinates")
            abi.check(L.tracer_set_connectivity(I(m.elem2D_nodes), I(m.edges), I(m.edge_tri),
                                                I(m.nod_in_elem2D), I(m.nod_in_elem2D_num), m.max_adj,
                                                m.nod2D, m.elem2D, m.edge2D), "tracer_set_connectivity")
            abi.check(L.tracer_set_levels(I(m.nlevels_nod2D), I(m.ulevels_nod2D), m.nod2D), "tracer_set_levels")
            abi.check(L.tracer_set_elem_levels(I(m.nlevels), I(m.ulevels), m.elem2D), "tracer_set_elem_levels")
            abi.check(L.tracer_set_geometry(D(m.elem_area), D(m.area), D(m.areasvol), D(m.edge_dxdy),
                                            D(m.edge_cross_dxdy), D(m.elem_cos), D(m.gradient_sca),
                                            m.nl, m.nod2D, m.elem2D, m.edge2D), "tracer_set_geometry")
            abi.check(L.tracer_set_vertical_grid(D(m.hnode), D(m.hnode_new), D(m.helem), D(m.Z_3d_n),
                                                 D(m.zbar_3d_n), m.nl, m.nod2D, m.elem2D),
                      "tracer_set_vertical_grid")
            abi.check(L.tracer_set_muscl_support(I(m.edge_up_dn_tri), I(m.nboundary_lay), m.nod2D,
                                                 m.edge2D), "tracer_set_muscl_support")
        abi.check(L.tracer_set_velocity_elem(D(case.uv), D(case.w), D(case.wi), D(case.we), m.nl,
                                             m.nod2D, m.elem2D), "tracer_set_velocity_elem")
        abi.check(L.tracer_set_wsplit(int(bool(getattr(case, "use_wsplit", False)))), "tracer_set_wsplit")
        dtn = getattr(case, "dt_node", None)
        abi.check(L.tracer_set_node_dt(None if dtn is None else D(dtn), m.nod2D), "tracer_set_node_dt")
        abi.check(L.tracer_allocate_tracers(case.ntracers, m.nl, m.nod2D, 2), "tracer_allocate_tracers")
        for i in range(case.ntracers):
            abi.check(L.tracer_set_values(i + 1, D(case.tracers[i]), m.nl, m.nod2D), "tracer_set_values")
            sc = case.schemes[i]
            abi.check(L.tracer_set_scheme(i + 1, sc.hor.encode(), sc.ver.encode(), sc.lim.encode(),
                                          float(sc.ph), float(sc.pv)), "tracer_set_scheme")
        if upload:
            abi.check(L.tracer_init_complete(), "tracer_init_complete")
        self._keep = []

    # --- execution options ------------------------------------------------------------------------
    def set_precision(self, mode: int) -> None:
        """0 FP64, 1 FP32, 2 FP32 arithmetic + FP16 storage, 3 FP32 arithmetic + BF16 storage."""
        abi.check(self.lib.tracer_set_precision(mode), "tracer_set_precision")

    def set_compensated(self, enable: bool) -> None:
        """Single precision: keep each stored value's rounding remainder (tracer_set_compensated)."""
        abi.check(self.lib.tracer_set_compensated(int(bool(enable))), "tracer_set_compensated")

    def set_rounding(self, stochastic: bool) -> None:
        """Stochastic (default) or round-to-nearest storage of the update in modes 1-3."""
        abi.check(self.lib.tracer_set_rounding(int(bool(stochastic))), "tracer_set_rounding")

    def set_block_width(self, kb: int) -> None:
        """Tracers interleaved per block and advanced per kernel pass (power of two, 0 = auto)."""
        abi.check(self.lib.tracer_set_tracers_per_pass(kb), "tracer_set_tracers_per_pass")

    def block_width(self) -> int:
        kb = ctypes.c_int()
        abi.check(self.lib.tracer_get_tracers_per_pass(ctypes.byref(kb)), "tracer_get_tracers_per_pass")
        return kb.value

    def set_fused(self, enable: bool) -> None:
        """Single fused kernel per step for FP64 FCT tracers in blocks of one."""
        abi.check(self.lib.tracer_set_fused(int(bool(enable))), "tracer_set_fused")

    def set_wsplit(self, enable: bool) -> None:
        """dynamics%use_wsplit: implicit low-order vertical part with w_i (FCT tracers)."""
        abi.check(self.lib.tracer_set_wsplit(int(bool(enable))), "tracer_set_wsplit")

    def wsplit(self) -> bool:
        v = ctypes.c_int()
        abi.check(self.lib.tracer_get_wsplit(ctypes.byref(v)), "tracer_get_wsplit")
        return bool(v.value)

    def set_ab_mode(self, mode: int, epsilon: float = 0.1) -> None:
        """valuesAB per step: 0 := values, 1 frozen (the reference's drivers), 2 FESOM2's AB2."""
        abi.check(self.lib.tracer_set_ab_mode(int(mode), float(epsilon)), "tracer_set_ab_mode")

    def ab_mode(self):
        mode, eps = ctypes.c_int(), ctypes.c_double()
        abi.check(self.lib.tracer_get_ab_mode(ctypes.byref(mode), ctypes.byref(eps)), "tracer_get_ab_mode")
        return mode.value, eps.value

    def set_halo_tiles(self, enable: bool) -> None:
        """LDS halo tiles for the FCT node kernels of blocks of width <= 4 (default on)."""
        abi.check(self.lib.tracer_set_halo_tiles(int(bool(enable))), "tracer_set_halo_tiles")

    def halo_tiles(self) -> int:
        """Number of halo tiles in use (0: untiled kernels)."""
        n = ctypes.c_int()
        abi.check(self.lib.tracer_get_halo_tiles(ctypes.byref(n)), "tracer_get_halo_tiles")
        return n.value

    def set_graph(self, enable: bool) -> None:
        abi.check(self.lib.tracer_set_graph(int(bool(enable))), "tracer_set_graph")

    def set_launch_config(self, lc) -> None:
        abi.check(self.lib.tracer_set_launch_config(*[int(x) for x in lc]), "tracer_set_launch_config")

    def launch_config(self):
        cfg = (ctypes.c_int * 4)()
        abi.check(self.lib.tracer_get_launch_config(cfg), "tracer_get_launch_config")
        return tuple(cfg)

    def kernel_times(self):
        """Mean ms per pass of the four stages since tracer_set_kernel_timing(1), and the count."""
        ms = (ctypes.c_double * 4)()
        cnt = ctypes.c_long()
        abi.check(self.lib.tracer_get_kernel_times(ms, ctypes.byref(cnt)), "tracer_get_kernel_times")
        return list(ms), cnt.value

    def step_times(self):
        """Mean ms per step of the edge-transport kernel and of the whole step (k_qflux through
        the last pass) since tracer_set_kernel_timing(1), and the number of steps."""
        q, st, cnt = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        abi.check(self.lib.tracer_get_step_times(ctypes.byref(q), ctypes.byref(st), ctypes.byref(cnt)),
                  "tracer_get_step_times")
        return q.value, st.value, cnt.value

    # --- stepping ---------------------------------------------------------------------------------
    def advect_steps(self, nsteps: int, dt: float) -> float:
        """Advance every tracer nsteps times; returns the device time in ms."""
        abi.check(self.lib.tracer_advect_steps(nsteps, dt), "tracer_advect_steps")
        ms = ctypes.c_double()
        self.lib.tracer_last_elapsed_ms(ctypes.byref(ms))
        return ms.value

    def do_oce_adv_tra(self, dt: float, tracer_id: int) -> None:
        abi.check(self.lib.tracer_do_oce_adv_tra(dt, tracer_id), "tracer_do_oce_adv_tra")

    def set_velocity(self, uv: np.ndarray, w: np.ndarray, wi: Optional[np.ndarray] = None,
                     we: Optional[np.ndarray] = None) -> None:
        """Element velocities uv (elem2D, nl-1, 2) and w, w_i, w_e (nod2D, nl); the device copy is
        updated in place (tracers and captured graphs are kept)."""
        arrs = [_f64(uv), _f64(w)] + [None if a is None else _f64(a) for a in (wi, we)]
        ptr = [None if a is None else abi.dptr(a) for a in arrs]
        abi.check(self.lib.tracer_set_velocity_elem(ptr[0], ptr[1], ptr[2], ptr[3], self.nl, self.nod2D,
                                                    self.elem2D), "tracer_set_velocity_elem")

    # --- ensembles ---------------------------------------------------------------------------------
    def set_ensemble(self, nmembers: int) -> None:
        """Number of velocity members on the mesh (tracer_set_ensemble)."""
        abi.check(self.lib.tracer_set_ensemble(int(nmembers)), "tracer_set_ensemble")

    def ensemble(self) -> int:
        n = ctypes.c_int()
        abi.check(self.lib.tracer_get_ensemble(ctypes.byref(n)), "tracer_get_ensemble")
        return n.value

    def set_velocity_member(self, member: int, uv: np.ndarray, w: np.ndarray,
                            wi: Optional[np.ndarray] = None, we: Optional[np.ndarray] = None) -> None:
        """Velocity of ensemble member `member` (1-based), layouts as set_velocity."""
        arrs = [_f64(uv), _f64(w)] + [None if a is None else _f64(a) for a in (wi, we)]
        ptr = [None if a is None else abi.dptr(a) for a in arrs]
        abi.check(self.lib.tracer_set_velocity_member(int(member), ptr[0], ptr[1], ptr[2], ptr[3], self.nl,
                                                      self.nod2D, self.elem2D), "tracer_set_velocity_member")

    def set_tracer_member(self, tracer_id: int, member: int) -> None:
        abi.check(self.lib.tracer_set_tracer_member(int(tracer_id), int(member)), "tracer_set_tracer_member")

    def setup_ensemble(self, cases, order=None) -> None:
        """Members that share a mesh and a time step, each with its own velocity and tracers
        (configs.ensemble_cases).  The library's tracer j + 1 is slot order[j] of the member-major
        list of (member, tracer) pairs (default: member by member); ``self.member_tracers[k][i]``
        is the library tracer id of tracer i of member k."""
        import dataclasses
        base = cases[0]
        for c in cases[1:]:
            if c.mesh is not base.mesh or c.dt != base.dt:
                raise ValueError("ensemble members share one mesh object and one time step")
        slots = [(k, i) for k, c in enumerate(cases) for i in range(c.ntracers)]
        if order is not None:
            slots = [slots[j] for j in order]
        tracers = np.stack([cases[k].tracers[i] for k, i in slots])
        schemes = [cases[k].schemes[i] for k, i in slots]
        allc = dataclasses.replace(base, tracers=tracers, schemes=schemes)
        self.setup_case(allc, upload=False)
        self.set_ensemble(len(cases))
        for k, c in enumerate(cases[1:], start=2):
            self.set_velocity_member(k, c.uv, c.w, c.wi, c.we)
        self.member_tracers = [[0] * c.ntracers for c in cases]
        for tid, (k, i) in enumerate(slots, start=1):
            self.set_tracer_member(tid, k + 1)
            self.member_tracers[k][i] = tid
        abi.check(self.lib.tracer_init_complete(), "tracer_init_complete")

    def checkpoint(self, restore: bool = False) -> None:
        """Keep (restore=False) or put back (restore=True) the tracers and their time levels."""
        abi.check(self.lib.tracer_checkpoint(int(bool(restore))), "tracer_checkpoint")

    def zero_tendencies(self) -> None:
        abi.check(self.lib.tracer_zero_tendencies(), "tracer_zero_tendencies")

    def set_values(self, tracer_id: int, values: np.ndarray) -> None:
        v = _f64(values)
        abi.check(self.lib.tracer_set_values(tracer_id, abi.dptr(v), self.nl, self.nod2D), "tracer_set_values")

    def set_values_ab(self, tracer_id: int, values: np.ndarray) -> None:
        v = _f64(values)
        abi.check(self.lib.tracer_set_values_ab(tracer_id, abi.dptr(v), self.nl, self.nod2D),
                  "tracer_set_values_ab")

    def get_tendencies(self):
        h = np.zeros((self.nod2D, self.nl - 1))
        v = np.zeros((self.nod2D, self.nl - 1))
        abi.check(self.lib.tracer_get_tendencies(abi.dptr(h), abi.dptr(v), self.nl - 1, self.nod2D),
                  "tracer_get_tendencies")
        return h, v

    def set_concurrent(self, enable: bool) -> None:
        """Low-order kernels on a second stream beside the edge-flux kernels (default True)."""
        abi.check(self.lib.tracer_set_concurrent(int(bool(enable))), "tracer_set_concurrent")

    def set_fused_limiter(self, enable: bool) -> None:
        """Limiter bounds, ratios and update in one kernel where it applies (default True)."""
        abi.check(self.lib.tracer_set_fused_limiter(int(bool(enable))), "tracer_set_fused_limiter")

    def fused_limiter_active(self) -> bool:
        v = ctypes.c_int(0)
        abi.check(self.lib.tracer_get_fused_limiter(ctypes.byref(v)), "tracer_get_fused_limiter")
        return bool(v.value)

    def set_keep_work(self, keep: bool) -> None:
        """Whether do_oce_adv_tra keeps the limited fluxes for get_work (default True)."""
        abi.check(self.lib.tracer_set_keep_work(int(bool(keep))), "tracer_set_keep_work")

    def get_work(self, name: str) -> np.ndarray:
        shapes = {"fct_LO": (self.nod2D, self.nl - 1), "adv_flux_hor": (self.edge2D, self.nl - 1),
                  "adv_flux_ver": (self.nod2D, self.nl)}
        out = np.zeros(shapes[name])
        abi.check(self.lib.tracer_get_work(name.encode(), abi.dptr(out), out.size), "tracer_get_work")
        return out

    def get_all_values(self) -> np.ndarray:
        return np.stack([self.get_tracer_values(i + 1) for i in range(self.num_tracers)])

    def stats(self, tracer_id: int) -> Dict[str, float]:
        a, b, c, d = (ctypes.c_double() for _ in range(4))
        abi.check(self.lib.tracer_get_stats_by_id(tracer_id, ctypes.byref(a), ctypes.byref(b),
                                                  ctypes.byref(c), ctypes.byref(d)),
                  "tracer_get_stats_by_id")
        return {"min": a.value, "max": b.value, "sum": c.value, "volume_sum": d.value}

    # --- PyTorch device tensors -------------------------------------------------------------------
    def set_values_tensor(self, tracer_id: int, t) -> None:
        """Copy a float64 CUDA/HIP tensor of shape (nod2D, nl-1) (reference numbering) in."""
        import torch
        assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        assert tuple(t.shape) == (self.nod2D, self.nl - 1)
        torch.cuda.current_stream().synchronize()
        abi.check(self.lib.tracer_set_values_device(tracer_id, ctypes.c_void_p(t.data_ptr()),
                                                    self.nl, self.nod2D), "tracer_set_values_device")

    def set_values_ab_tensor(self, tracer_id: int, t) -> None:
        """valuesAB from a float64 device tensor of shape (nod2D, nl-1) (reference numbering)."""
        import torch
        assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        assert tuple(t.shape) == (self.nod2D, self.nl - 1)
        torch.cuda.current_stream().synchronize()
        abi.check(self.lib.tracer_set_values_ab_device(tracer_id, ctypes.c_void_p(t.data_ptr()),
                                                       self.nl, self.nod2D), "tracer_set_values_ab_device")

    def get_tendencies_tensor(self, dttf_h=None, dttf_v=None):
        """del_ttf_advhoriz / advvert into float64 device tensors (nod2D, nl-1)."""
        import torch
        shape = (self.nod2D, self.nl - 1)
        if dttf_h is None:
            dttf_h = torch.empty(shape, dtype=torch.float64, device="cuda")
        if dttf_v is None:
            dttf_v = torch.empty(shape, dtype=torch.float64, device="cuda")
        for t in (dttf_h, dttf_v):
            assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and tuple(t.shape) == shape
        torch.cuda.current_stream().synchronize()
        abi.check(self.lib.tracer_get_tendencies_device(ctypes.c_void_p(dttf_h.data_ptr()),
                                                        ctypes.c_void_p(dttf_v.data_ptr()),
                                                        self.nl - 1, self.nod2D), "tracer_get_tendencies_device")
        return dttf_h, dttf_v

    def get_values_tensor(self, tracer_id: int, out=None):
        import torch
        if out is None:
            out = torch.empty((self.nod2D, self.nl - 1), dtype=torch.float64, device="cuda")
        torch.cuda.current_stream().synchronize()
        abi.check(self.lib.tracer_get_values_device(tracer_id, ctypes.c_void_p(out.data_ptr()),
                                                    self.nl - 1, self.nod2D), "tracer_get_values_device")
        return out

    def cleanup(self) -> None:
        self.lib.tracer_cleanup()
