"""Host driver over the C-ABI, mirroring the reference's module API.

Method names follow the reference procedures they stand in for; run_model
restates the iteration of src/main.f90:124-227 (transport -> rates -> cell
update -> opacity feedback until converged, then one final consistency pass).
"""
import ctypes as C
import time

import numpy as np

from .lib import DUST_FIELDS, DUSTRT_KEYS, FIELDS, MC_DUSTRT_NREPORT, MC_GRID_AMR, Params, check, library


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class Model:
    """One device-resident model: grid, leaf state, tallies, rates."""

    def __init__(self, params=None, **kw):
        L = library()
        self.params = params if isinstance(params, Params) else Params.make(**(params or {}), **kw)
        self._h = C.c_void_p()
        check(L.mc_init(C.byref(self.params), C.byref(self._h)), "mc_init")
        self._L = L
        self._refresh_info()
        if self.params.diffuse_field and self.params.hei_diffuse:
            self.diffuse_hei_setup()       # src/main.f90:107-112

    def _refresh_info(self):
        n, nnu, lt = C.c_int32(), C.c_int32(), C.c_double()
        check(self._L.mc_info(self._h, C.byref(n), C.byref(nnu), C.byref(lt), None), "mc_info")
        self.nleaf, self.nnu_band, self.ion_Ltot = n.value, nnu.value, lt.value

    # ---- set-up --------------------------------------------------------------
    def grid_octree(self, xleaf, yleaf, zleaf, level, half):
        """amr_build_tree + amr_build_neighbors from a leaf list."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (xleaf, yleaf, zleaf))
        lev = np.ascontiguousarray(level, dtype=np.int32)
        check(self._L.mc_grid_octree(self._h, lev.size, _dp(x), _dp(y), _dp(z),
                                     lev.ctypes.data_as(C.POINTER(C.c_int32)), float(half)),
              "mc_grid_octree")
        self.leaves = (x, y, z, lev, float(half))
        self._refresh_info()

    def gas_state_setup(self, nH, xHI=None):
        nH = np.ascontiguousarray(nH, dtype=np.float64)
        if nH.size != self.nleaf:
            raise ValueError(f"nH has {nH.size} values, grid has {self.nleaf} leaves")
        x = None if xHI is None else np.ascontiguousarray(xHI, dtype=np.float64)
        check(self._L.mc_gas_state_setup(self._h, _dp(nH), None if x is None else _dp(x)),
              "mc_gas_state_setup")

    def species_setup(self, elements):
        """species_setup: load the trace-metal registry (list of element dicts
        from mochii_amd.atomic.registry, with 'abund' set)."""
        from .atomic import to_struct
        self.elements = list(elements)
        self._els = to_struct(self.elements)
        check(self._L.mc_species_setup(self._h, len(self.elements), self._els),
              "mc_species_setup")

    def metal_segment_transitions(self):
        """Metal transitions summed over the last pass's scored segments."""
        v = C.c_int64()
        check(self._L.mc_metal_segment_transitions(self._h, C.byref(v)),
              "mc_metal_segment_transitions")
        return v.value

    def walk_atomics(self):
        """Global FP64 tally atomics the walk issued in the last pass."""
        v = C.c_int64()
        check(self._L.mc_walk_atomics(self._h, C.byref(v)), "mc_walk_atomics")
        return v.value

    def species_fractions(self, ie):
        """stage_frac of registry element ie: (nleaf, nstage)."""
        ns = self.elements[ie]["nstage"]
        out = np.zeros((self.nleaf, ns))
        check(self._L.mc_species_fractions_get(self._h, ie, _dp(out)), "mc_species_fractions_get")
        return out

    def species_rates(self, ie):
        """(egam, eheat) of registry element ie: (nleaf, nstage-1) each."""
        nt = self.elements[ie]["nstage"] - 1
        g = np.zeros((self.nleaf, nt))
        h = np.zeros((self.nleaf, nt))
        check(self._L.mc_species_rates_get(self._h, ie, _dp(g), _dp(h)), "mc_species_rates_get")
        return g, h

    def species_tallies(self, ie):
        """raw(2, nstage-1) of registry element ie per leaf: (nleaf, nstage-1, 2)."""
        nt = self.elements[ie]["nstage"] - 1
        out = np.zeros((self.nleaf, nt, 2))
        check(self._L.mc_species_tallies_get(self._h, ie, _dp(out)), "mc_species_tallies_get")
        return out

    # ---- dust thermal-infrared band -----------------------------------------
    def dust_temp_setup(self):
        """dust_temp_setup: femit(T) of the loaded extinction table."""
        check(self._L.mc_dust_temp_setup(self._h), "mc_dust_temp_setup")
        self.dust_temp_ready = True

    def dust_temp_compute(self):
        """dust_temp_compute: T_dust of every leaf from the current Heat_dust."""
        check(self._L.mc_dust_temp_compute(self._h), "mc_dust_temp_compute")
        return self.dust("t_dust")

    def dust_emis_transport_setup(self):
        check(self._L.mc_dust_emis_transport_setup(self._h), "mc_dust_emis_transport_setup")

    def transport_dust_emission(self):
        """transport_dust_emission: the infrared self-heating iteration on the
        device-resident Heat_dust; returns its report as a dict."""
        rep = np.zeros(MC_DUSTRT_NREPORT)
        check(self._L.mc_transport_dust_emission(self._h, _dp(rep)), "mc_transport_dust_emission")
        return dict(zip(DUSTRT_KEYS, [float(v) for v in rep]))

    def dust(self, name):
        """Per-leaf dust field: heat_dust, t_dust, heat_ir or ir_abs_path."""
        out = np.empty(self.nleaf)
        check(self._L.mc_dust_get(self._h, DUST_FIELDS[name], _dp(out)), "mc_dust_get")
        return out

    def set_heat_dust(self, heat):
        a = np.ascontiguousarray(heat, dtype=np.float64)
        check(self._L.mc_dust_set(self._h, DUST_FIELDS["heat_dust"], _dp(a)), "mc_dust_set")

    def dust_ir_energy_budget(self):
        out = np.zeros(3)
        check(self._L.mc_dust_ir_energy_budget(self._h, _dp(out)), "mc_dust_ir_energy_budget")
        return out

    def dust_ir_escaped_spectrum(self):
        nl = self.dust_ir_grid()["nl"]
        lam, Llam, dlam = np.zeros(nl), np.zeros(nl), np.zeros(nl)
        check(self._L.mc_dust_ir_escaped_spectrum(self._h, _dp(lam), _dp(Llam), _dp(dlam)),
              "mc_dust_ir_escaped_spectrum")
        return lam, Llam, dlam

    def dust_ir_grid(self):
        nl, nz, lh = C.c_int32(), C.c_int32(), C.c_double()
        check(self._L.mc_dust_ir_grid(self._h, C.byref(nl), None, None, None, None, None, None),
              "mc_dust_ir_grid")
        n = nl.value
        lam, sabs, edge, fe = np.zeros(n), np.zeros(n), np.zeros(n + 1), np.zeros(122)
        check(self._L.mc_dust_ir_grid(self._h, C.byref(nl), _dp(lam), _dp(sabs), _dp(edge),
                                      C.byref(nz), C.byref(lh), _dp(fe)), "mc_dust_ir_grid")
        return dict(nl=n, lam=lam, sabs=sabs, edge=edge, nzero=nz.value, lam_hardest=lh.value,
                    tgrid=fe[:61].copy(), femit=fe[61:].copy())

    def dust_ir_timing(self):
        t, sg = np.zeros(3), np.zeros(2, dtype=np.int64)
        check(self._L.mc_dust_ir_timing(self._h, _dp(t), sg.ctypes.data_as(C.POINTER(C.c_int64))),
              "mc_dust_ir_timing")
        return dict(launch_ms=t[0], walk_ms=t[1], other_ms=t[2], seg_direct=int(sg[0]),
                    seg_scattered=int(sg[1]))

    def ion_dust_table(self, lam, albedo, gcos, cext):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (lam, albedo, gcos, cext)]
        check(self._L.mc_ion_dust_table(self._h, arrs[0].size, *[_dp(a) for a in arrs]),
              "mc_ion_dust_table")

    def set_state(self, **fields):
        for k, v in fields.items():
            a = np.ascontiguousarray(v, dtype=np.float64)
            check(self._L.mc_state_set(self._h, FIELDS[k], _dp(a)), "mc_state_set")

    def state(self, *names):
        names = names or ("nH", "xHI", "xHeI", "xHeII", "ne", "Te")
        out = {}
        for k in names:
            a = np.empty(self.nleaf)
            check(self._L.mc_state_get(self._h, FIELDS[k], _dp(a)), "mc_state_get")
            out[k] = a
        return out

    def set_packet_order(self, order):
        """MC_ORDER_DIRECTION (default) or MC_ORDER_LAUNCH walk order."""
        check(self._L.mc_set_packet_order(self._h, int(order)), "mc_set_packet_order")

    def set_launch_batch(self, max_records):
        """Launch records per batch of a pass (default 2^30 mixed, 2^26 FP64)."""
        check(self._L.mc_set_launch_batch(self._h, int(max_records)), "mc_set_launch_batch")

    def set_row_issue(self, mode):
        """Issue form of the mixed walk's moment rows: MC_ROWS_AUTO (default),
        MC_ROWS_LANE or MC_ROWS_TOGETHER."""
        check(self._L.mc_set_row_issue(self._h, int(mode)), "mc_set_row_issue")

    def row_issue(self):
        """(mode, form of the last pass: 1 transposed, 0 per lane, -1 none)."""
        mode, last = C.c_int(), C.c_int()
        check(self._L.mc_row_issue(self._h, C.byref(mode), C.byref(last)), "mc_row_issue")
        return mode.value, last.value

    def set_precision(self, precision):
        check(self._L.mc_set_precision(self._h, int(precision)), "mc_set_precision")
        self.params.precision = int(precision)

    # ---- the hot path ----------------------------------------------------------
    def ionizing_field_and_rates(self, with_peel=False):
        check(self._L.mc_ionizing_field_and_rates(self._h, int(bool(with_peel))),
              "mc_ionizing_field_and_rates")

    def diffuse_info(self):
        """(diffuse_lum [erg/s], diffuse_nphot) of the last pass."""
        L, n = C.c_double(), C.c_int64()
        check(self._L.mc_diffuse_info(self._h, C.byref(L), C.byref(n)), "mc_diffuse_info")
        return L.value, n.value

    def diffuse_hei_setup(self, y=None, A=None):
        """hei_2ph_setup: the He I 2^1S two-photon shape (default: the data
        snapshot's HeI2phot.dat rows) for channel 4 of the diffuse field."""
        if y is None or A is None:
            from .atomic import load_hei_two_photon
            y, A = load_hei_two_photon()
        y = np.ascontiguousarray(y, dtype=np.float64)
        A = np.ascontiguousarray(A, dtype=np.float64)
        if y.shape != A.shape:
            raise ValueError("diffuse_hei_setup: y and A differ in length")
        check(self._L.mc_diffuse_hei_setup(self._h, y.size, _dp(y), _dp(A)),
              "mc_diffuse_hei_setup")

    def diffuse_hei_info(self):
        """(photons per decay, eV per decay, grid_y[1024], cdf_e[1024]) of the
        two-photon window."""
        g = np.zeros(1024)
        c = np.zeros(1024)
        n, e = C.c_double(), C.c_double()
        check(self._L.mc_diffuse_hei_info(self._h, C.byref(n), C.byref(e), _dp(g), _dp(c)),
              "mc_diffuse_hei_info")
        return n.value, e.value, g, c

    def gas_equilibrium_update(self):
        a, b = C.c_double(), C.c_double()
        check(self._L.mc_gas_equilibrium_update(self._h, C.byref(a), C.byref(b)),
              "mc_gas_equilibrium_update")
        return a.value, b.value

    def cooling_setup(self, fit):
        """fit: sequence of (A_i, T_i) terms of the H I line-cooling fit."""
        A = np.ascontiguousarray([f[0] for f in fit], dtype=np.float64)
        Ti = np.ascontiguousarray([f[1] for f in fit], dtype=np.float64)
        check(self._L.mc_cooling_setup(self._h, A.size, _dp(A), _dp(Ti)), "mc_cooling_setup")

    def nlevel_setup(self, atoms=None):
        """The n-level atoms ({'<el>_<stage>': atom}, default: every atom of
        the data snapshot): the (T, n_e) cooling tables of cooling_model
        'local_ne' and the lines of lines_compute."""
        from .atomic import load_nlevel, nlevel_struct
        self.atoms = dict(load_nlevel() if atoms is None else atoms)
        self._nla, n = nlevel_struct(self.atoms)
        check(self._L.mc_nlevel_setup(self._h, n, self._nla), "mc_nlevel_setup")

    def lines_setup(self, line_case="B", case_ab=None):
        """sh95_setup: the H I / He II recombination-line tables of the line
        case (par%line_case 'A', 'B' or 'follow' = the continuum case
        case_ab) and the He I (Porter, case B) table of the data snapshot."""
        from .atomic import LineTable, load_line_tables
        tabs = load_line_tables()
        if case_ab is None:
            case_ab = "A" if self.params.case_A else "B"
        case = case_ab if line_case == "follow" else line_case
        self.line_case = case
        self._line_src = [tabs[f"sh95_hi_case{case}"], tabs.get(f"sh95_heii_case{case}"),
                          tabs.get("hei_porter_caseB")]
        self._lts = [LineTable.make(t) if t is not None else None for t in self._line_src]
        args = [C.byref(t) if t is not None else None for t in self._lts]
        check(self._L.mc_lines_setup(self._h, *args), "mc_lines_setup")

    def lines_compute(self):
        """lines_write's luminosities on the current state: dict of arrays
        L [erg/s], wl [A], ion (100 Z + stage) and labels (recombination
        line names, '' for the n-level rows), in the reference's order."""
        n = C.c_int32()
        check(self._L.mc_lines_compute(self._h, 0, None, None, None, C.byref(n)),
              "mc_lines_compute")
        L, wl = np.zeros(n.value), np.zeros(n.value)
        ion = np.zeros(n.value, dtype=np.int32)
        check(self._L.mc_lines_compute(self._h, n.value, _dp(L), _dp(wl),
                                       ion.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)),
              "mc_lines_compute")
        labels = [ln["label"] for t in self._line_src if t is not None for ln in t["lines"]]
        labels += [""] * (n.value - len(labels))
        return dict(L=L, wl=wl, ion=ion, labels=labels)

    def gas_thermal_update(self):
        a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        check(self._L.mc_gas_thermal_update(self._h, C.byref(a), C.byref(b), C.byref(c),
                                            C.byref(d)), "mc_gas_thermal_update")
        return a.value, b.value, c.value, d.value

    # ---- learned warm start (no counterpart in the reference) -----------------
    def set_tally_moments(self, on):
        """Energy-moment tallies of the MC_MIXED metal walk (default on)."""
        check(self._L.mc_set_tally_moments(self._h, int(bool(on))), "mc_set_tally_moments")

    def tally_moment_info(self):
        """(sub-bins over the band, largest relative interpolation error);
        (0, 0.0) when the walk keeps the per-coefficient sums."""
        n, e = C.c_int32(), C.c_double()
        check(self._L.mc_tally_moment_info(self._h, C.byref(n), C.byref(e)),
              "mc_tally_moment_info")
        return n.value, e.value

    def set_warm_start(self, mode):
        check(self._L.mc_set_warm_start(self._h, int(mode)), "mc_set_warm_start")

    def set_warm_levels(self, levels):
        """Bisection levels a warm-started thermal solve runs first."""
        check(self._L.mc_set_warm_levels(self._h, int(levels)), "mc_set_warm_levels")

    def qmc_uniforms(self, stream, first_index, n, ndim):
        """The launch's Sobol/Owen uniforms of stream 1 (stellar) or 2
        (diffuse), computed on the device: (n, ndim)."""
        out = np.zeros((n, ndim))
        check(self._L.mc_qmc_uniforms(self._h, int(stream), int(first_index), int(n), int(ndim),
                                      _dp(out)), "mc_qmc_uniforms")
        return out

    def launch_records(self, first_id, n):
        """Stellar launch records of ids first_id .. first_id+n-1 as the
        pass computes them: (rec (n, 7) x y z kx ky kz E, bin_leaf (n, 2))."""
        rec = np.zeros((n, 7))
        bl = np.zeros((n, 2), dtype=np.int32)
        check(self._L.mc_launch_records(self._h, int(first_id), int(n), _dp(rec),
                                        bl.ctypes.data_as(C.POINTER(C.c_int32))),
              "mc_launch_records")
        return rec, bl

    def thermal_stats(self):
        a = np.zeros(4, dtype=np.int64)
        check(self._L.mc_thermal_stats(self._h, a.ctypes.data_as(C.POINTER(C.c_int64))),
              "mc_thermal_stats")
        return dict(evaluations=int(a[0]), fallbacks=int(a[1]), bracketed=int(a[2]),
                    leaves=int(a[3]))

    def mlp_set(self, params):
        p = _f32(params)
        check(self._L.mc_mlp_set(self._h, _fp(p), p.size), "mc_mlp_set")

    def mlp_get(self):
        from .lib import MC_MLP_NPARAM
        p = np.zeros(MC_MLP_NPARAM, dtype=np.float32)
        check(self._L.mc_mlp_get(self._h, _fp(p)), "mc_mlp_get")
        return p

    def mlp_features(self):
        X = np.zeros((self.nleaf, 8), dtype=np.float32)
        check(self._L.mc_mlp_features(self._h, _fp(X)), "mc_mlp_features")
        return X

    def mlp_train(self, X, Y, steps, batch=1024, lr=1e-3, seed=1):
        X, Y = _f32(X), _f32(Y)
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        check(self._L.mc_mlp_train(self._h, _fp(X), _fp(Y), X.shape[0], int(steps), int(batch),
                                   float(lr), int(seed), _fp(hist)), "mc_mlp_train")
        return hist[:steps]

    def mlp_grad(self, X, Y):
        from .lib import MC_MLP_NPARAM
        X, Y = _f32(X), _f32(Y)
        g = np.zeros(MC_MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        check(self._L.mc_mlp_grad(self._h, _fp(X), _fp(Y), X.shape[0], C.byref(loss), _fp(g)),
              "mc_mlp_grad")
        return loss.value, g

    def mlp_infer(self, X):
        X = _f32(X)
        out = np.zeros((X.shape[0], 4), dtype=np.float32)
        check(self._L.mc_mlp_infer(self._h, _fp(X), X.shape[0], _fp(out)), "mc_mlp_infer")
        return out

    def hydrogen_balance_residual_grid(self):
        a, b = C.c_double(), C.c_double()
        check(self._L.mc_hydrogen_balance_residual_grid(self._h, C.byref(a), C.byref(b)),
              "mc_hydrogen_balance_residual_grid")
        return a.value, b.value

    def gas_opacity_fill(self):
        check(self._L.mc_gas_opacity_fill(self._h), "mc_gas_opacity_fill")

    # ---- results -----------------------------------------------------------------
    def tallies(self):
        n, nnu = self.nleaf, self.nnu_band
        hhe, jt, dust = np.empty((n, 6)), np.empty((n, nnu)), np.empty(n)
        segs = np.zeros(2, dtype=np.int64)
        check(self._L.mc_tallies_get(self._h, _dp(hhe), _dp(jt), _dp(dust),
                                     segs.ctypes.data_as(C.POINTER(C.c_int64))), "mc_tallies_get")
        return dict(hhe=hhe, jt=jt, dust_abs=dust, seg_direct=int(segs[0]),
                    seg_scattered=int(segs[1]))

    def segment_counts(self):
        """{direct, scattered} scored segments of the last pass, without
        copying the tallies."""
        segs = np.zeros(2, dtype=np.int64)
        check(self._L.mc_tallies_get(self._h, None, None, None,
                                     segs.ctypes.data_as(C.POINTER(C.c_int64))), "mc_tallies_get")
        return int(segs[0]), int(segs[1])

    def rates(self):
        r = np.empty(6 * self.nleaf)
        check(self._L.mc_rates_get(self._h, _dp(r)), "mc_rates_get")
        r = r.reshape(6, self.nleaf)
        return dict(gHI=r[0], gHeI=r[1], gHeII=r[2], hHI=r[3], hHeI=r[4], hHeII=r[5])

    def energy_ledger(self):
        a = np.empty(3)
        check(self._L.mc_energy_ledger(self._h, _dp(a)), "mc_energy_ledger")
        return a

    def band_edges(self):
        e = np.empty(self.nnu_band + 1)
        check(self._L.mc_info(self._h, None, None, None, _dp(e)), "mc_info")
        return e

    def timing(self):
        t = np.empty(4)
        check(self._L.mc_timing(self._h, _dp(t)), "mc_timing")
        return dict(launch_ms=t[0], walk_ms=t[1], rates_ms=t[2], solve_ms=t[3])

    def trace_ray(self, pos, direction, precision=None, maxn=100000):
        prec = self.params.precision if precision is None else precision
        p = np.ascontiguousarray(pos, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        cells = np.zeros(maxn, dtype=np.int32)
        n = C.c_int32()
        check(self._L.mc_trace_ray(self._h, int(prec), _dp(p), _dp(d),
                                   cells.ctypes.data_as(C.POINTER(C.c_int32)), maxn, C.byref(n)),
              "mc_trace_ray")
        return cells[:n.value].copy()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mc_finalize(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_model(model, niter=None, log=None, final_pass=True):
    """The iteration of src/main.f90:124-227 on a set-up Model.

    Returns a dict with the iteration count, convergence flag, the metrics of
    the last update and wall-clock times."""
    p = model.params
    niter = max(p.gas_niter, 1) if niter is None else niter
    converged = False
    hist = []
    t0 = time.perf_counter()
    it = 0
    for it in range(1, niter + 1):
        model.ionizing_field_and_rates()
        if p.gas_niter < 1:
            break
        if p.solve_te:
            max_dx, max_dte, dx_vol, dte_vol = model.gas_thermal_update()
            if p.conv_vol:
                converged = dx_vol < p.gas_tol and dte_vol < p.gas_tol_te
            else:
                converged = max_dx < p.gas_tol and max_dte < p.gas_tol_te
            hist.append((it, max_dx, dx_vol, max_dte, dte_vol))
        else:
            max_dx, dx_vol = model.gas_equilibrium_update()
            converged = (dx_vol < p.gas_tol) if p.conv_vol else (max_dx < p.gas_tol)
            hist.append((it, max_dx, dx_vol))
        if log:
            log(f"iteration {it:4d}: " + ", ".join(f"{v:.4e}" for v in hist[-1][1:]))
        if converged:
            break
    t_loop = time.perf_counter() - t0
    if final_pass and p.gas_niter >= 1:
        model.ionizing_field_and_rates()
    return dict(iterations=it, converged=converged, history=hist, t_loop=t_loop,
                t_total=time.perf_counter() - t0)
