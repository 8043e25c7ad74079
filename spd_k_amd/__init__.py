"""spd_k_amd — MI355X spectral-difference update (hydro + induction) for spd_K.

Python mirror of the library's C ABI (include/spdk.h). The kernels are the
gfx950 HIP kernels in libspdk.so; PyTorch only provides device memory and the
stream. Every call fails loudly when the library or a gfx950 device is
missing: there is no CPU fallback.

Two levels, as in the reference:
  * `run(args)`: the whole driver (src/main.cpp), an argv list of
    "-i file.athinput" and "block/name=value" overrides; outputs use the
    reference's raw float64 file format.
  * `Context`: one method per kernel-layer function of src/prototypes.hpp
    (transform_a_to_b_1d, compute_fluxes, sd_riemann_solver, boundaries,
    update_prediction, update_solution, ...) on CUDA tensors in the
    reference's C-order layouts.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libspdk.so")
DRIVER_PATH = os.path.join(_HERE, "_build", "spdk")

SPDK_M_SP_TO_FP, SPDK_M_FP_TO_SP, SPDK_M_DFP_TO_SP, SPDK_M_SP_TO_CV, SPDK_M_CV_TO_SP, SPDK_M_FP_TO_CV, SPDK_M_INVADER = range(7)
BC = {"periodic": 0, "gradfree": 1, "reflective": 2}


class SDShape(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nt", "nv", "Nz", "Ny", "Nx", "nz", "ny", "nx")]

    @classmethod
    def of(cls, shape):
        return cls(*[int(x) for x in shape])


class FVShape(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nv", "Nz", "Ny", "Nx")]

    @classmethod
    def of(cls, shape):
        return cls(*[int(x) for x in shape])


class Params(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("active", ctypes.c_int * 3), ("bc", ctypes.c_int * 3),
                ("n_ader", ctypes.c_int), ("rk_order", ctypes.c_int),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double), ("g", ctypes.c_double * 3),
                ("nu", ctypes.c_double), ("h", ctypes.c_double * 3), ("nad_tol", ctypes.c_double),
                ("nad_delta", ctypes.c_int), ("nad_moore", ctypes.c_int), ("sed", ctypes.c_int),
                ("blending", ctypes.c_int), ("vel", ctypes.c_double * 3), ("beta", ctypes.c_double)]


class Operators(ctypes.Structure):
    _fields_ = [("mat", ctypes.c_void_p * 7), ("wt", ctypes.c_void_p), ("fv_faces", ctypes.c_void_p * 3),
                ("fv_centers", ctypes.c_void_p * 3), ("fv_nfaces", ctypes.c_int * 3), ("fv_ncells", ctypes.c_int * 3)]


class RunResult(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("dofs", ctypes.c_longlong), ("evolution_seconds", ctypes.c_double),
                ("dof_updates_per_s", ctypes.c_double), ("t_final", ctypes.c_double),
                ("kernel_ms_total", ctypes.c_double), ("kernel_launches", ctypes.c_int),
                ("kernel_name", ctypes.c_char * 64), ("picard_steps", ctypes.c_longlong),
                ("picard_sweeps", ctypes.c_double), ("picard_fallbacks", ctypes.c_longlong),
                ("picard_hist", ctypes.c_longlong * 32)]


_lib = None


def lib():
    """The loaded libspdk.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python spd_k_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        L.spdk_last_error.restype = ctypes.c_char_p
        L.spdk_fused_workspace_bytes.restype = ctypes.c_size_t
        L.spdk_fused_fluxes_workspace_bytes.restype = ctypes.c_size_t
        L.spdk_fused_noalloc_max_bytes.restype = ctypes.c_size_t
        L.spdk_ctx_stream.restype = ctypes.c_void_p
        L.spdk_mlp_param_count.restype = ctypes.c_size_t
        for f in ("spdk_mlp_set_params", "spdk_mlp_get_params", "spdk_mlp_destroy", "spdk_mlp_dims"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.spdk_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.spdk_mlp_param_count.argtypes = [ctypes.c_void_p]
        L.spdk_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_long]
        L.spdk_mlp_gradients.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                                         ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.spdk_mlp_train_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                                          ctypes.c_long, ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
        _lib = L
    return _lib


def last_error():
    return lib().spdk_last_error().decode(errors="replace")


def check(rc, what="spdk call"):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def run(args, outdir="output"):
    """Runs the host driver in-process; returns the run result as a dict."""
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    res = RunResult()
    rc = lib().spdk_run_argv(len(args), argv, outdir.encode(), ctypes.byref(res))
    if rc != 0:
        raise RuntimeError(f"spdk_run_argv failed ({rc}): {last_error()}")
    return {"steps": res.steps, "dofs": res.dofs, "evolution_seconds": res.evolution_seconds,
            "dof_updates_per_s": res.dof_updates_per_s, "t_final": res.t_final,
            "kernel_ms_total": res.kernel_ms_total, "kernel_launches": res.kernel_launches,
            "kernel_name": res.kernel_name.decode(), "picard_steps": res.picard_steps,
            "picard_sweeps": res.picard_sweeps, "picard_fallbacks": res.picard_fallbacks,
            "picard_hist": list(res.picard_hist)}


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("spd_k_amd kernels take contiguous device tensors")
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """Kernel-layer interface (src/prototypes.hpp) over libspdk.so.

    `ops` is a dict of float64 numpy arrays: the seven operator matrices
    (keys sp_to_fp, fp_to_sp, dfp_to_sp, sp_to_cv, cv_to_sp, fp_to_cv,
    invader), `wt`, and per axis `fv_faces`/`fv_centers` lists.
    """

    MATS = ("sp_to_fp", "fp_to_sp", "dfp_to_sp", "sp_to_cv", "cv_to_sp", "fp_to_cv", "invader")

    def __init__(self, p, active, bc, n_ader, ops, gamma=1.4, cfl=0.8, h=(1.0, 1.0, 1.0), g=(0.0, 0.0, 0.0), nu=0.0,
                 nad_tol=1e-5, nad_delta=False, nad_moore=True, sed=True, blending=True, rk_order=0,
                 vel=(1.0, 1.0, 0.0), beta=None):
        import numpy as np
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("spd_k_amd.Context needs a GPU (no device visible)")
        self._keep = []
        prm = Params()
        prm.p = p
        for d in range(3):
            prm.active[d] = int(active[d])
            prm.bc[d] = int(bc[d]) if not isinstance(bc[d], str) else BC[bc[d]]
            prm.g[d] = g[d]
            prm.h[d] = h[d]
            prm.vel[d] = vel[d]
        prm.n_ader, prm.rk_order = n_ader, rk_order
        prm.gamma, prm.cfl, prm.nu, prm.nad_tol = gamma, cfl, nu, nad_tol
        prm.beta = -2.0 / 3.0 * nu if beta is None else beta  # hydro/beta default (src/main.cpp:236-237)
        prm.nad_delta, prm.nad_moore, prm.sed, prm.blending = int(nad_delta), int(nad_moore), int(sed), int(blending)
        o = Operators()

        def hold(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data

        for q, name in enumerate(self.MATS):
            o.mat[q] = hold(ops[name])
        o.wt = hold(ops["wt"])
        for d in range(3):
            f = ops.get("fv_faces", [None] * 3)[d]
            c = ops.get("fv_centers", [None] * 3)[d]
            if f is not None:
                o.fv_faces[d] = hold(f)
                o.fv_nfaces[d] = len(f)
            if c is not None:
                o.fv_centers[d] = hold(c)
                o.fv_ncells[d] = len(c)
        self.ctx = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().spdk_ctx_create(ctypes.byref(prm), ctypes.byref(o), ctypes.c_void_p(stream), ctypes.byref(self.ctx)),
              "spdk_ctx_create")

    def close(self):
        if self.ctx:
            lib().spdk_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().spdk_ctx_sync(self.ctx), "spdk_ctx_sync")

    # --- src/transforms.cpp
    def transform_a_to_b_1d(self, A, B, mat, dim, t_src=-1):
        check(lib().spdk_transform_1d(self.ctx, _ptr(A), SDShape.of(A.shape), _ptr(B), SDShape.of(B.shape), mat, dim, t_src),
              "transform_1d")

    def transform_a_to_b(self, A, B, T, mat):
        check(lib().spdk_transform(self.ctx, _ptr(A), _ptr(B), _ptr(T), SDShape.of(A.shape), mat), "transform")

    def update_prediction(self, U, U_ader, Fx, Fy, Fz, q, dt):
        check(lib().spdk_ader_predict(self.ctx, _ptr(U), _ptr(U_ader), _ptr(Fx), _ptr(Fy), _ptr(Fz),
                                      SDShape.of(U_ader.shape), q, ctypes.c_double(dt)), "ader_predict")

    def update_solution(self, U, U_ader, Fx, Fy, Fz, q, dt):
        check(lib().spdk_ader_correct(self.ctx, _ptr(U), _ptr(U_ader), _ptr(Fx), _ptr(Fy), _ptr(Fz),
                                      SDShape.of(U_ader.shape), q, ctypes.c_double(dt)), "ader_correct")

    def combine_solution(self, U, U0, a):
        check(lib().spdk_rk_combine(self.ctx, _ptr(U), _ptr(U0), SDShape.of(U.shape), ctypes.c_double(a)), "rk_combine")

    def copy_ader(self, U, U_ader):
        check(lib().spdk_copy_ader(self.ctx, _ptr(U), _ptr(U_ader), SDShape.of(U_ader.shape)), "copy_ader")

    # --- src/hydro.cpp
    def compute_conservatives(self, W, U):
        check(lib().spdk_conservatives(self.ctx, _ptr(W), _ptr(U), SDShape.of(W.shape)), "conservatives")

    def compute_primitives(self, U, W):
        check(lib().spdk_primitives(self.ctx, _ptr(U), _ptr(W), SDShape.of(U.shape)), "primitives")

    def compute_fluxes(self, U, F, dim):
        check(lib().spdk_hydro_flux(self.ctx, _ptr(U), _ptr(F), SDShape.of(U.shape), dim), "hydro_flux")

    def sd_riemann_solver(self, U, F, dim, viscous=False):
        check(lib().spdk_sd_riemann(self.ctx, _ptr(U), _ptr(F), SDShape.of(U.shape), dim, int(viscous)), "sd_riemann")

    # --- viscous terms (src/hydro.cpp:420-555)
    def compute_gradient(self, U_fp, dU, dim):
        check(lib().spdk_sd_gradient(self.ctx, _ptr(U_fp), SDShape.of(U_fp.shape), _ptr(dU), SDShape.of(dU.shape), dim),
              "sd_gradient")

    def compute_viscous_flux(self, U_fp, dU1, dU2, dU3, dim):
        check(lib().spdk_viscous_flux(self.ctx, _ptr(U_fp), SDShape.of(U_fp.shape), _ptr(dU1), _ptr(dU2), _ptr(dU3),
                                      SDShape.of(dU1.shape), dim), "viscous_flux")

    def sd_rusanov_solver(self, U_fp, F, dim):
        check(lib().spdk_sd_rusanov(self.ctx, _ptr(U_fp), _ptr(F), SDShape.of(U_fp.shape), dim), "sd_rusanov")

    def compute_dt(self, W_cv):
        out = ctypes.c_double()
        check(lib().spdk_min_dt(self.ctx, _ptr(W_cv), SDShape.of(W_cv.shape), ctypes.byref(out)), "min_dt")
        return out.value

    # --- src/boundary.cpp
    def boundaries(self, U, dim):
        check(lib().spdk_fill_sd_traces(self.ctx, _ptr(U), SDShape.of(U.shape), dim), "fill_sd_traces")

    def fv_boundaries(self, U):
        check(lib().spdk_fill_fv_ghosts(self.ctx, _ptr(U), FVShape.of(U.shape)), "fill_fv_ghosts")

    # --- FV fallback (src/transforms.cpp, src/trouble_detection.cpp, src/hydro.cpp)
    def face_integral(self, F_fp, F, T, t_id, dim):
        check(lib().spdk_fv_face_integral(self.ctx, _ptr(F_fp), SDShape.of(F_fp.shape), _ptr(F), FVShape.of(F.shape),
                                          _ptr(T), t_id, dim), "fv_face_integral")

    def fv_update_solution(self, U_new, U_old, U_cv, Fx, Fy, Fz, t_id, dt, update):
        check(lib().spdk_fv_update(self.ctx, _ptr(U_new), _ptr(U_old), _ptr(U_cv), SDShape.of(U_cv.shape), _ptr(Fx),
                                   _ptr(Fy), _ptr(Fz), FVShape.of(U_new.shape), t_id, ctypes.c_double(dt), int(update)),
              "fv_update")

    def fv_update_prims(self, U_new, W_new, W_old, U_cv, Fx, Fy, Fz, t_id, dt):
        check(lib().spdk_fv_update_prims(self.ctx, _ptr(U_new), _ptr(W_new), _ptr(W_old), _ptr(U_cv),
                                         SDShape.of(U_cv.shape), _ptr(Fx), _ptr(Fy), _ptr(Fz), FVShape.of(U_new.shape),
                                         t_id, ctypes.c_double(dt)), "fv_update_prims")

    def fv_update_blended(self, U_new, U_cv, Fx, Fy, Fz, theta, t_id, dt):
        check(lib().spdk_fv_update_blended(self.ctx, _ptr(U_new), _ptr(U_cv), SDShape.of(U_cv.shape), _ptr(Fx), _ptr(Fy),
                                           _ptr(Fz), _ptr(theta), FVShape.of(U_new.shape), t_id, ctypes.c_double(dt)),
              "fv_update_blended")

    def compute_primitives_fv(self, U, W):
        check(lib().spdk_fv_primitives(self.ctx, _ptr(U), _ptr(W), FVShape.of(U.shape)), "fv_primitives")

    def detect_troubles(self, W_new, W_old, troubles, ax, ay, az, pad, mask):
        check(lib().spdk_detect_troubles(self.ctx, _ptr(W_new), _ptr(W_old), _ptr(troubles), _ptr(ax), _ptr(ay), _ptr(az),
                                         FVShape.of(W_new.shape), int(pad), int(mask)), "detect_troubles")

    def detect_troubles_bits(self, W_new, W_old, troubles, abits, pad, mask):
        """detect_troubles with the SED alphas as 16-bit words per cell
        (abits: int16 tensor of the FV cell count, zero before the first call)."""
        check(lib().spdk_detect_troubles_bits(self.ctx, _ptr(W_new), _ptr(W_old), _ptr(troubles), _ptr(abits),
                                              FVShape.of(W_new.shape), int(pad), int(mask)), "detect_troubles_bits")

    def blend_theta(self, troubles, theta_tmp, theta):
        check(lib().spdk_blend_theta(self.ctx, _ptr(troubles), _ptr(theta_tmp), _ptr(theta), FVShape.of(troubles.shape)),
              "blend_theta")

    # --- Solve_fluxes (src/hydro_ader.hpp:278-289) with the fused tile kernel
    def solve_fluxes_fused(self, U_ader, Fx, Fy, Fz):
        """Flux-point fluxes of every slice of U_ader (sp->fp, fluxes, trace
        fill, LLF) into Fx/Fy/Fz (None for an inactive axis)."""
        import torch
        L = lib()
        N = (ctypes.c_int * 3)(U_ader.shape[4], U_ader.shape[3], U_ader.shape[2])
        nb = L.spdk_fused_fluxes_workspace_bytes(self.ctx, N)
        if nb == 0:
            raise RuntimeError(f"fused_fluxes: {last_error()}")
        ws = torch.empty(nb, dtype=torch.uint8, device=U_ader.device)
        check(L.spdk_fused_fluxes(self.ctx, _ptr(U_ader), SDShape.of(U_ader.shape), _ptr(ws), _ptr(Fx), _ptr(Fy),
                                  _ptr(Fz)), "fused_fluxes")
        self.sync()

    def fallback_fluxes(self, W, theta, Fx, Fy, Fz, dt_sub):
        check(lib().spdk_muscl_hllc_blend(self.ctx, _ptr(W), _ptr(theta), _ptr(Fx), _ptr(Fy), _ptr(Fz),
                                          FVShape.of(W.shape), ctypes.c_double(dt_sub)), "muscl_hllc_blend")

    # --- one whole step (src/hydro_ader.hpp:291-325 + compute_dt) with the fused tile kernel
    def hydro_step_fused(self, U_sp, dt, W_cv=None):
        """One plain FP64 step on U_sp with a workspace allocated here; returns
        the next dt (the call blocks for it)."""
        import torch
        L = lib()
        N = (ctypes.c_int * 3)(U_sp.shape[4], U_sp.shape[3], U_sp.shape[2])
        nb = L.spdk_fused_workspace_bytes(self.ctx, N)
        if nb == 0:
            raise RuntimeError(f"fused step: {last_error()}")
        ws = torch.empty(nb, dtype=torch.uint8, device=U_sp.device)
        out = ctypes.c_double()
        check(L.spdk_hydro_step_fused(self.ctx, _ptr(U_sp), SDShape.of(U_sp.shape), _ptr(ws), ctypes.c_double(dt),
                                      _ptr(W_cv), None, ctypes.byref(out)), "hydro_step_fused")
        return out.value

    def hydro_step_fused_noalloc(self, U_sp, dt, W_cv=None):
        """The same step with the library's static per-GPU workspace (no
        allocation in the call; at most spdk_fused_noalloc_max_bytes() of
        workspace; calls overlapping on one GPU race)."""
        out = ctypes.c_double()
        check(lib().spdk_hydro_step_fused_noalloc(self.ctx, _ptr(U_sp), SDShape.of(U_sp.shape), ctypes.c_double(dt),
                                                  _ptr(W_cv), None, ctypes.byref(out)), "hydro_step_fused_noalloc")
        return out.value
