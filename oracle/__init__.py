"""CPU oracle for the MPMWarpYDW per-substep hot path (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker or the timed CPU baseline.
The product path (``mpmwarpydw_amd``) never imports it.

``Oracle(64)`` wraps ``liboracle_f64.so`` (FP64, the oracle of record) and
``Oracle(32)`` wraps ``liboracle_f32.so`` (FP32 restatement of the reference's
arithmetic).  Every method takes numpy arrays in the reference's AoS layout
(vec3 ``(N,3)``, mat33 ``(N,3,3)``, grid ``(gx,gy,gz[,3])``) and mutates the
output arrays in place, exactly as the reference kernels mutate Warp arrays.
Each wrapped C function cites the reference kernel it restates
(utils/mpmRoutines.py, utils/xpbdRoutines.py, utils/simulationRoutines.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_I = C.c_int
_L = C.c_long
_D = C.c_double
_P = C.c_void_p


def build() -> None:
    """Compile both oracle precisions and the OpenMP FP32 build with the
    Makefile in this folder."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """ctypes view of one precision of the C oracle."""

    def __init__(self, bits: int = 64, svd_from_ftf: bool | None = None, omp: bool = False):
        """``omp``: the FP32 build with OpenMP loops (the all-cores CPU baseline)."""
        if omp and bits != 32:
            raise ValueError("the OpenMP build is FP32")
        path = os.path.join(HERE, f"liboracle_f{bits}{'_omp' if omp else ''}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.bits = bits
        self.dtype = np.float64 if bits == 64 else np.float32
        self.R = C.c_double if bits == 64 else C.c_float
        assert self.lib.o_real_size() == np.dtype(self.dtype).itemsize
        if svd_from_ftf is None:
            svd_from_ftf = bits == 32  # the FP32 restatement resolves F^T F like an SVD of F
        self.lib.o_set_svd_mode(int(svd_from_ftf))
        self.lib.o_velocity_convergence.restype = C.c_double
        self.lib.o_sum_active.restype = C.c_double

    def threads(self) -> int:
        """Threads of the OpenMP build's loops (1 for the serial builds)."""
        return int(self.lib.o_threads())

    # ---- helpers ---------------------------------------------------------
    def f(self, a):
        """Return ``a`` as a C-contiguous array of the oracle's REAL type (copy if needed)."""
        return np.ascontiguousarray(a, dtype=self.dtype)

    @staticmethod
    def i(a):
        return np.ascontiguousarray(a, dtype=np.int32)

    def r(self, x):
        return self.R(float(x))

    def _chk(self, *arrs):
        for a in arrs:
            if a is None:
                continue
            assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
            assert a.dtype in (self.dtype, np.int32), a.dtype

    # ---- A5 svd3 ---------------------------------------------------------
    def svd3(self, F):
        F = self.f(F).reshape(3, 3)
        U = np.zeros((3, 3), self.dtype)
        V = np.zeros((3, 3), self.dtype)
        s = np.zeros(3, self.dtype)
        lam = np.zeros(3, self.dtype)
        self.lib.o_svd3(_ptr(F), _ptr(U), _ptr(s), _ptr(V), _ptr(lam))
        return U, s, V

    # ---- A1 --------------------------------------------------------------
    def compute_mu_lam_bulk(self, E, nu, mu, lam, bulk):
        self._chk(E, nu, mu, lam, bulk)
        self.lib.o_compute_mu_lam_bulk(_I(len(E)), _ptr(E), _ptr(nu), _ptr(mu), _ptr(lam), _ptr(bulk))

    def initialize_geostatic_F(self, x, F, mu, lam, density, g, z_top, K0, plane_strain=False):
        """initialize_geostatic_F (mpmRoutines.py:52-124) or its plane-strain form (:128-203)."""
        self._chk(x, F, mu, lam, density)
        self.lib.o_initialize_geostatic_F(_I(len(mu)), _ptr(x), _ptr(F), _ptr(mu), _ptr(lam), _ptr(density),
                                          self.r(g), self.r(z_top), self.r(K0), _I(int(bool(plane_strain))))

    def deactivate_and_teleport(self, x, v, activeLabel, materialLabel, z_datum, teleport):
        """deactivate_and_teleport_xpbd_below_datum (mpmRoutines.py:1446-1475); labels in place."""
        self._chk(x, v, activeLabel, materialLabel)
        self.lib.o_deactivate_and_teleport(_I(len(activeLabel)), _ptr(x), _ptr(v), _ptr(activeLabel),
                                           _ptr(materialLabel), self.r(z_datum), self.r(teleport))

    # ---- A2/A3/A4 --------------------------------------------------------
    def compute_stress_from_F_trial(self, activeLabel, materialLabel, x, v, x_init, v_init, F, F_trial,
                                    mu, lam, ys, hardening, softening, density, strainCriteria, eff,
                                    eta_shear, eta_bulk, C_, stress, acc, damage, constitutive_model, alpha):
        """Argument order of mpmRoutines.compute_stress_from_F_trial (mpmRoutines.py:229-254)."""
        self._chk(activeLabel, materialLabel, x, v, x_init, v_init, F, F_trial, mu, lam, ys, hardening,
                  softening, density, strainCriteria, eta_shear, eta_bulk, C_, stress, acc, damage, alpha)
        self.lib.o_compute_stress_from_F_trial(
            _I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel), _ptr(x), _ptr(v), _ptr(x_init),
            _ptr(v_init), _ptr(F), _ptr(F_trial), _ptr(mu), _ptr(lam), _ptr(ys), _ptr(hardening),
            _ptr(softening), _ptr(density), _ptr(strainCriteria), self.r(eff), _ptr(eta_shear),
            _ptr(eta_bulk), _ptr(C_), _ptr(stress), _ptr(acc), _ptr(damage), _I(int(constitutive_model)),
            _ptr(alpha))

    # ---- A6 --------------------------------------------------------------
    def p2g_apic_with_stress(self, activeLabel, materialLabel, stress, x, v, C_, vol, mass, dx, inv_dx,
                             minBounds, rpic_damping, grid_m, grid_v_in, dt):
        """Argument order of mpmRoutines.p2g_apic_with_stress (mpmRoutines.py:485-501)."""
        self._chk(activeLabel, materialLabel, stress, x, v, C_, vol, mass, grid_m, grid_v_in)
        gx, gy, gz = grid_m.shape
        mb = self.f(minBounds)
        return self.lib.o_p2g(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel), _ptr(stress),
                              _ptr(x), _ptr(v), _ptr(C_), _ptr(vol), _ptr(mass), self.r(dx), self.r(inv_dx),
                              _ptr(mb), self.r(rpic_damping), _I(gx), _I(gy), _I(gz), _ptr(grid_m),
                              _ptr(grid_v_in), self.r(dt))

    # ---- A7/A8/A9 --------------------------------------------------------
    def grid_normalization_and_gravity(self, grid_m, grid_v_in, grid_v_out, gravity, dt):
        self._chk(grid_m, grid_v_in, grid_v_out)
        g = self.f(gravity)
        self.lib.o_grid_normalization_and_gravity(_L(grid_m.size), _ptr(grid_m), _ptr(grid_v_in),
                                                  _ptr(grid_v_out), _ptr(g), self.r(dt))

    def add_damping_via_grid(self, grid_v_out, scale):
        self._chk(grid_v_out)
        self.lib.o_add_damping(_L(grid_v_out.size // 3), _ptr(grid_v_out), self.r(scale))

    def collideBounds(self, grid_v_out, gx, gy, gz, friction, dt, gmag, padding):
        self._chk(grid_v_out)
        self.lib.o_collide_bounds(_ptr(grid_v_out), _I(gx), _I(gy), _I(gz), self.r(friction), self.r(dt),
                                  self.r(gmag), _I(padding))

    def collide_bounds_variant(self, kind, grid_v_out, gx, gy, gz, friction, dt, gmag, padding, restitution=0.0):
        """The grid boundary conditions other than "friction" (simulationRoutines.py:141-228)."""
        self._chk(grid_v_out)
        v = _ptr(grid_v_out)
        if kind == "friction_gradual":
            self.lib.o_collide_bounds_gradual(v, _I(gx), _I(gy), _I(gz), self.r(friction), self.r(dt), self.r(gmag),
                                              _I(padding))
        elif kind == "restitution":
            self.lib.o_collide_bounds_restitution(v, _I(gx), _I(gy), _I(gz), self.r(restitution), self.r(friction),
                                                  self.r(dt), self.r(gmag), _I(padding))
        elif kind == "absorbing":
            self.lib.o_collide_bounds_absorbing(v, _I(gx), _I(gy), _I(gz), _I(padding))
        elif kind == "plane_strain_y":
            self.lib.o_collide_bounds_plane_strain_y(v, _I(gx), _I(gy), _I(gz), self.r(friction), self.r(dt),
                                                     self.r(gmag), _I(padding))
        elif kind == "plane_strain_y_global":
            self.lib.o_collide_bounds_plane_strain_y_global(v, _I(gx), _I(gy), _I(gz), self.r(friction),
                                                            self.r(dt), self.r(gmag), _I(padding))
        else:
            raise ValueError(kind)

    def volumetric_locking_correction(self, activeLabel, materialLabel, x, C_, vol, inv_dx, minBounds, gx, gy, gz):
        """Both VLC phases (mpmRoutines.py:805-911); returns (cell_div, cell_vol) and updates C in place."""
        self._chk(x, C_, vol)
        G = gx * gy * gz
        cd = np.zeros(G, self.dtype)
        cv = np.zeros(G, self.dtype)
        mb = self.f(minBounds)
        self.lib.o_vlc(_I(len(vol)), _ptr(self.i(activeLabel)), _ptr(self.i(materialLabel)), _ptr(x), _ptr(C_),
                       _ptr(vol), self.r(inv_dx), _ptr(mb), _I(gx), _I(gy), _I(gz), _ptr(cd), _ptr(cv))
        return cd.reshape(gx, gy, gz), cv.reshape(gx, gy, gz)

    # ---- A10 -------------------------------------------------------------
    def g2p(self, dt, activeLabel, materialLabel, x, v, C_, F, F_trial, cov, inv_dx, grid_v_out,
            update_cov, minBounds):
        """Argument order of mpmRoutines.g2p (mpmRoutines.py:715-728)."""
        self._chk(activeLabel, materialLabel, x, v, C_, F, F_trial, cov, grid_v_out)
        gx, gy, gz = grid_v_out.shape[:3]
        mb = self.f(minBounds)
        return self.lib.o_g2p(self.r(dt), _I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel),
                              _ptr(x), _ptr(v), _ptr(C_), _ptr(F), _ptr(F_trial), _ptr(cov), self.r(inv_dx),
                              _ptr(grid_v_out), _I(gx), _I(gy), _I(gz), _I(int(bool(update_cov))), _ptr(mb))

    # ---- XPBD ------------------------------------------------------------
    def reset_mpm_contact_labels(self, activeLabel, materialLabel):
        self._chk(activeLabel, materialLabel)
        self.lib.o_reset_mpm_contact_labels(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel))

    def integrateParticlesXPBD(self, activeLabel, materialLabel, x, v, gravity, dt, x_new, v_new, v_max):
        self._chk(activeLabel, materialLabel, x, v, x_new, v_new)
        g = self.f(gravity)
        self.lib.o_integrate_particles_xpbd(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel),
                                            _ptr(x), _ptr(v), _ptr(g), self.r(dt), _ptr(x_new), _ptr(v_new),
                                            self.r(v_max))

    def my_solve_particle_bound_contacts(self, activeLabel, materialLabel, x, v, mass, radius, mu, thr,
                                         static_mu, minb, maxb, dt, relaxation, delta):
        self._chk(activeLabel, materialLabel, x, v, mass, radius, delta)
        self.lib.o_solve_particle_bound_contacts(
            _I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel), _ptr(x), _ptr(v), _ptr(mass),
            _ptr(radius), self.r(mu), self.r(thr), self.r(static_mu), _ptr(self.f(minb)), _ptr(self.f(maxb)),
            self.r(dt), self.r(relaxation), _ptr(delta))

    def hashgrid_build(self, x, cell_width, dims=(128, 128, 128)):
        """wp.HashGrid(dims).build(x, cell_width) -> (point_ids, cell_starts, cell_ends)."""
        x = self.f(x)
        n = len(x)
        ncell = int(np.prod(dims))
        ids = np.zeros(n, np.int32)
        cs = np.zeros(ncell, np.int32)
        ce = np.zeros(ncell, np.int32)
        self.lib.o_hashgrid_build(_I(n), _ptr(x), self.r(cell_width), _I(dims[0]), _I(dims[1]), _I(dims[2]),
                                  _ptr(ids), _ptr(cs), _ptr(ce))
        return ids, cs, ce

    def hashgrid_query(self, grid, pos, radius):
        """Candidate ids of wp.hash_grid_query(grid, pos, radius) in visiting order."""
        ids, cs, ce, width, dims = grid
        out = np.zeros(len(ids) + 1, np.int32)
        k = self.lib.o_hashgrid_query(_ptr(self.f(pos)), self.r(radius), _ptr(ids), _ptr(cs), _ptr(ce),
                                      self.r(width), _I(dims[0]), _I(dims[1]), _I(dims[2]), _ptr(out),
                                      _I(len(out)))
        return out[:k].tolist()

    def my_solve_particle_particle_contacts(self, activeLabel, materialLabel, grid, x, v, mass, radius, k_mu,
                                            thr, static_mu, k_cohesion, max_radius, dt, relaxation, lock,
                                            deltas, order=0):
        """``grid`` = (point_ids, cell_starts, cell_ends, cell_width, dims).
        ``order`` 0 is wp.hash_grid_query's candidate order (x fastest, ids
        ascending); 1 is the reverse and 2 has z fastest (for bounding what the
        unpinned order changes)."""
        ids, cs, ce, width, dims = grid
        self._chk(activeLabel, materialLabel, x, v, mass, radius, deltas)
        self.lib.o_solve_particle_particle_contacts_ordered(
            _I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel), _ptr(ids), _ptr(cs), _ptr(ce),
            self.r(width), _I(dims[0]), _I(dims[1]), _I(dims[2]), _ptr(x), _ptr(v), _ptr(mass), _ptr(radius),
            self.r(k_mu), self.r(thr), self.r(static_mu), self.r(k_cohesion), self.r(max_radius), self.r(dt),
            self.r(relaxation), _I(int(lock)), _ptr(deltas), _I(int(order)))

    def my_apply_particle_deltas(self, activeLabel, materialLabel, x_orig, x_pred, delta, dt, v_max, x_out,
                                 v_out):
        self._chk(activeLabel, materialLabel, x_orig, x_pred, delta, x_out, v_out)
        self.lib.o_apply_particle_deltas(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel),
                                         _ptr(x_orig), _ptr(x_pred), _ptr(delta), self.r(dt), self.r(v_max),
                                         _ptr(x_out), _ptr(v_out))

    def apply_mpm_contact_impulses(self, activeLabel, materialLabel, x_orig, v_orig, x_final, v_final, dt,
                                   v_max, coupling, xpbd_only, v_out):
        self._chk(activeLabel, materialLabel, x_orig, v_orig, x_final, v_final, v_out)
        self.lib.o_apply_mpm_contact_impulses(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel),
                                              _ptr(x_orig), _ptr(v_orig), _ptr(x_final), _ptr(v_final),
                                              self.r(dt), self.r(v_max), self.r(coupling), _I(int(xpbd_only)),
                                              _ptr(v_out))

    def sleepParticles(self, activeLabel, materialLabel, thr, radius, x_prev, x_after, dist_total, dt):
        self._chk(activeLabel, materialLabel, radius, x_prev, x_after, dist_total)
        cnt = np.zeros(2, np.int32)
        self.lib.o_sleep_particles(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel), self.r(thr),
                                   _ptr(radius), _ptr(x_prev), _ptr(x_after), _ptr(dist_total), self.r(dt),
                                   _ptr(cnt))
        return int(cnt[0]), int(cnt[1])

    def swellParticlesType2(self, activeLabel, materialLabel, x_init, x, radius, factor, factor_max, swell_r,
                            base_r):
        self._chk(activeLabel, materialLabel, x_init, x, radius, swell_r, base_r)
        self.lib.o_swell_particles_type2(_I(len(activeLabel)), _ptr(activeLabel), _ptr(materialLabel),
                                         _ptr(x_init), _ptr(x), _ptr(radius), self.r(factor),
                                         self.r(factor_max), _ptr(swell_r), _ptr(base_r))

    # ---- A12 -------------------------------------------------------------
    def velocityConvergence(self, activeLabel, v, radius):
        self._chk(activeLabel, v, radius)
        return self.lib.o_velocity_convergence(_I(len(activeLabel)), _ptr(activeLabel), _ptr(v), _ptr(radius))

    def countActiveParticles(self, activeLabel):
        return int(self.lib.o_count_active(_I(len(activeLabel)), _ptr(self.i(activeLabel))))

    def countParticlesByType(self, activeLabel, materialLabel):
        out = np.zeros(4, np.int32)
        self.lib.o_count_by_type(_I(len(activeLabel)), _ptr(self.i(activeLabel)), _ptr(self.i(materialLabel)),
                                 _ptr(out))
        return tuple(int(v) for v in out)

    def sum_float_array_active(self, arr, activeLabel):
        cnt = C.c_int(0)
        s = self.lib.o_sum_active(_I(len(arr)), _ptr(self.f(arr)), _ptr(self.i(activeLabel)), C.byref(cnt))
        return s, cnt.value


class MLPOracle:
    """FP64 CPU oracle of the learned return map (oracle/mlp_oracle.c):
    the 10-64-64-3 GELU network's forward pass, mean-squared-error gradient and
    Adam step, with optional bfloat16 rounding at the points where the GPU
    kernels feed their matrix units (``quant=True``), plus the features,
    exact targets and judge of the learned projection."""

    NI, NH, NO = 10, 64, 3

    def __init__(self):
        path = os.path.join(HERE, "liboracle_mlp.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.lib.o_bf16.restype = C.c_double
        self.lib.o_bf16.argtypes = [C.c_double]
        self.lib.o_bf16_split.restype = C.c_double
        self.lib.o_bf16_split.argtypes = [C.c_double]
        self.lib.o_gelu.restype = C.c_double
        self.lib.o_gelu.argtypes = [C.c_double]
        self.lib.o_gelu_grad.restype = C.c_double
        self.lib.o_gelu_grad.argtypes = [C.c_double]
        self.lib.o_mlp_loss_grad.restype = C.c_double
        self.nparams = int(self.lib.o_mlp_nparams())

    @staticmethod
    def _d(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def bf16(self, v):
        return self.lib.o_bf16(float(v))

    def bf16_split(self, v):
        """An input as the kernels feed it: bf16(v) + bf16(v - bf16(v))."""
        return self.lib.o_bf16_split(float(v))

    def gelu(self, v):
        return self.lib.o_gelu(float(v))

    def gelu_grad(self, v):
        return self.lib.o_gelu_grad(float(v))

    def split(self, params):
        """Views W1 (64,10), b1, W2 (64,64), b2, W3 (3,64), b3 of a flat parameter vector."""
        p = np.asarray(params)
        o = 0
        out = []
        for shape in ((64, 10), (64,), (64, 64), (64,), (3, 64), (3,)):
            k = int(np.prod(shape))
            out.append(p[o:o + k].reshape(shape))
            o += k
        return out

    def forward(self, params, X, quant=False):
        X = self._d(X)
        params = self._d(params)
        Y = np.zeros((X.shape[0], 3))
        self.lib.o_mlp_forward(_I(X.shape[0]), _ptr(params), _ptr(X), _ptr(Y), _I(int(quant)))
        return Y

    def loss_grad(self, params, X, T, quant=False):
        X, T, params = self._d(X), self._d(T), self._d(params)
        g = np.zeros(self.nparams)
        loss = self.lib.o_mlp_loss_grad(_I(X.shape[0]), _ptr(params), _ptr(X), _ptr(T), _ptr(g), _I(int(quant)))
        return loss, g

    def adam_step(self, params, grad, m, v, lr, beta1, beta2, eps, t):
        """In place on params, m, v (float64 arrays)."""
        for a in (params, m, v):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        grad = self._d(grad)
        self.lib.o_adam_step(_I(len(params)), _ptr(params), _ptr(grad), _ptr(m), _ptr(v), _D(lr), _D(beta1),
                             _D(beta2), _D(eps), _I(t))

    def features(self, e, mu, lam, ys, alpha, acc, damage):
        n = len(mu)
        X = np.zeros((n, 10))
        self.lib.o_lrm_features(_I(n), *[_ptr(self._d(a)) for a in (e, mu, lam, ys, alpha, acc, damage)], _ptr(X))
        return X

    def targets(self, e, mu, lam, ys, alpha, damage):
        n = len(mu)
        T = np.zeros((n, 3))
        plastic = np.zeros(n, np.int32)
        self.lib.o_lrm_targets(_I(n), *[_ptr(self._d(a)) for a in (e, mu, lam, ys, alpha, damage)], _ptr(T),
                               _ptr(plastic))
        return T, plastic.astype(bool)

    def apply(self, e, Y, mu, lam, ys, alpha, damage):
        """The predicted projection e0' + eref Y / 10 (radial return plus the
        network's correction) of trial strains e from network outputs Y."""
        n = len(mu)
        out = np.zeros((n, 3))
        self.lib.o_lrm_apply(_I(n), *[_ptr(self._d(a)) for a in (e, mu, lam, ys, alpha, damage, Y)], _ptr(out))
        return out

    def judge(self, e, e_pred, mu, lam, ys, alpha, damage, tol=1e-4):
        n = len(mu)
        acc = np.zeros(n, np.int32)
        self.lib.o_lrm_judge(_I(n), *[_ptr(self._d(a)) for a in (e, e_pred, mu, lam, ys, alpha, damage)], _D(tol),
                             _ptr(acc))
        return acc.astype(bool)
