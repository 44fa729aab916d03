"""Survey-like estimators restated in numpy (oracle; test infrastructure only).

Follows, on full complex128 c2c grids exactly as the reference does:

* ``compute_particle_sums_pk`` / ``compute_particle_sums_bk``
  (``mesh_generator.py:1131-1268``),
* ``get_mesh_pk_survey`` (``mesh_generator.py:227-538``),
* ``get_mesh_bk_survey`` (``mesh_generator.py:541-1004``),
* ``get_N_field`` (``mesh_generator.py:1007-1128``),
* ``calculate_power_spectrum_survey`` + ``get_G_ell``
  (``clustering_estimator.py:205-563``, ``:3104-3134``),
* ``calculate_bk_sugi_survey`` (``clustering_estimator.py:1558-3101``).

Catalogs are passed as already-read structured arrays with the fields the
reference's ``catalog_reader`` produces (``Position``, ``WEIGHT``,
``WEIGHT_FKP``, ``NZ``; ``catalog_processor.py:531-563``); reading, z cuts
and RA/DEC/Z conversion are restated in ``cnpc_ref.catalog``.

The interlaced survey field is complex: the +1/2 k.H interlacing phase is not
Hermitian on the Nyquist planes, so ``c2r`` of the combined spectrum leaves an
imaginary part that the reference keeps (pmesh ``dtype="complex128"``) and
that the x-space Y_lm products carry into every k bin.  This restatement keeps
it exactly as the reference does.
"""
import numpy as np

from . import angular as A
from . import conventions as C


# ---------------------------------------------------------------- particle sums
def compute_particle_sums_pk(data, randoms, correlation_mode):
    """``mesh_generator.py:1131-1184``."""
    wd = data["WEIGHT"] * data["WEIGHT_FKP"]
    wr = randoms["WEIGHT"] * randoms["WEIGHT_FKP"]
    out = [
        np.sum(wd), np.sum(wr),
        np.sum(data["WEIGHT"]), np.sum(randoms["WEIGHT"]),
        np.sum(data["WEIGHT"] * data["NZ"] * data["WEIGHT_FKP"] * data["WEIGHT_FKP"]),
        np.sum(randoms["WEIGHT"] * randoms["NZ"] * randoms["WEIGHT_FKP"] * randoms["WEIGHT_FKP"]),
    ]
    if correlation_mode == "auto":
        out += [np.sum(wd ** 2), np.sum(wr ** 2)]
    else:
        out += [0.0, 0.0]
    return tuple(out)


def compute_particle_sums_bk(data, randoms, correlation_mode, angu_config):
    """``mesh_generator.py:1187-1268``: alpha sums, I_33 and S_LM[M], M=0..L,
    with Racah-normalised y_LM at the observer-frame unit position."""
    wd = data["WEIGHT"] * data["WEIGHT_FKP"]
    wr = randoms["WEIGHT"] * randoms["WEIGHT_FKP"]
    I_33 = np.sum(np.array(randoms["NZ"]) ** 2 * np.array(randoms["WEIGHT"]) ** 2
                  * np.array(randoms["WEIGHT_FKP"]) ** 3)
    if correlation_mode == "auto":
        dp, rp = data["Position"], randoms["Position"]
        du = dp / np.linalg.norm(dp, axis=1, keepdims=True)
        ru = rp / np.linalg.norm(rp, axis=1, keepdims=True)
        sd, sr = [], []
        L = angu_config[2]
        for mm in range(0, L + 1):
            y = A.get_Ylm(L, mm, Racah_normalized=True)
            sd.append(np.sum(wd ** 3 * np.conj(y(du[:, 0], du[:, 1], du[:, 2]))))
            sr.append(np.sum(wr ** 3 * np.conj(y(ru[:, 0], ru[:, 1], ru[:, 2]))))
        sd, sr = np.array(sd), np.array(sr)
    else:
        sd, sr = 0.0, 0.0
    return [np.sum(wd), np.sum(wr), np.sum(data["WEIGHT"]), np.sum(randoms["WEIGHT"]), I_33, sd, sr]


def survey_boxcenter(randoms):
    """Box centre from the first randoms catalog: ``0.5 * (min + max)`` of the
    positions (``catalog_processor.py:565-574``)."""
    p = randoms["Position"]
    return 0.5 * (np.min(p, axis=0) + np.max(p, axis=0))


def _alpha(sums, alpha_scheme):
    if alpha_scheme == "pypower":
        return sums[0] / sums[1]
    if alpha_scheme == "nbodykit":
        return sums[2] / sums[3]
    raise ValueError("alpha_scheme must be either 'pypower' or 'nbodykit'.")


def _paint(cat, weight, nmesh, boxsize, sampler, interlaced, boxcenter):
    return C.pm_painter(cat["Position"], weight, nmesh, boxsize, sampler, interlaced, boxcenter=boxcenter)


# ---------------------------------------------------------------- pk meshes
def get_mesh_pk_survey(cats, correlation_mode, nmesh, boxsize, sampler, interlaced,
                       normalization_scheme, alpha_scheme="pypower", boxcenter=None):
    """``mesh_generator.py:227-538`` with catalogs already read.

    Returns (mesh_attrs, rfield_a, rfield_b)."""
    n, L = int(nmesh), float(boxsize)
    vpc = L ** 3 / n ** 3
    attrs = {}
    if boxcenter is None:
        boxcenter = survey_boxcenter(cats["randoms_a"])
    attrs["boxcenter"] = boxcenter
    da, ra = cats["data_a"], cats["randoms_a"]
    sa = compute_particle_sums_pk(da, ra, correlation_mode)
    alpha_a = _alpha(sa, alpha_scheme)
    I_rand = alpha_a * sa[5]
    I_data = sa[4]
    attrs.update({"alpha_a": alpha_a, "I_rand": I_rand, "I_data": I_data})
    rd_a = _paint(da, da["WEIGHT"] * da["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
    rr_a = _paint(ra, ra["WEIGHT"] * ra["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
    rd_a = rd_a / vpc
    rr_a = rr_a / vpc
    rfield_a = rd_a - alpha_a * rr_a
    if correlation_mode == "auto":
        rfield_b = rfield_a
    else:
        db, rb = cats["data_b"], cats["randoms_b"]
        sb = compute_particle_sums_pk(db, rb, correlation_mode)
        alpha_b = _alpha(sb, alpha_scheme)
        attrs["alpha_b"] = alpha_b
        rd_b = _paint(db, db["WEIGHT"] * db["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter) / vpc
        rr_b = _paint(rb, rb["WEIGHT"] * rb["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter) / vpc
        rfield_b = rd_b - alpha_b * rr_b
    if normalization_scheme == "particle":
        if correlation_mode == "auto":
            attrs["N0"] = (sa[6] + alpha_a ** 2 * sa[7]) / I_rand
        else:
            raise ValueError("Particle-normalization scheme will not work for cross-correlation!")
    else:
        if correlation_mode == "auto":
            I_mesh = alpha_a * np.sum(rd_a * rr_a) * L ** 3 / n ** 3
        else:
            I_mesh = 0.5 * L ** 3 / n ** 3 * (alpha_b * np.sum(rd_a * rr_b) + alpha_a * np.sum(rd_b * rr_a))
        I_mesh = np.real(I_mesh).astype(np.float64)
        attrs["I_mesh"] = I_mesh
        attrs["N0"] = (sa[6] + alpha_a ** 2 * sa[7]) / I_mesh if correlation_mode == "auto" else 0.0
    return attrs, rfield_a, rfield_b


# ---------------------------------------------------------------- pk estimator
def get_G_ell(rfield, ell, kgrid, xgrid, comp, w, boxsize):
    """``clustering_estimator.py:3104-3134``: non-Racah Y_lm, m = 0..ell."""
    ylms = [A.get_Ylm(ell, m) for m in range(ell + 1)]
    G = C.r2c(rfield * ylms[0](*xgrid)) * ylms[0](*kgrid)
    for m in range(1, ell + 1):
        cf = C.r2c(rfield * np.conj(ylms[m](*xgrid))) * ylms[m](*kgrid)
        G = G + 2 * cf
    G = comp(w, G)
    return G * (4 * np.pi * float(boxsize) ** 3 / (2 * ell + 1))


def space_inversion(G):
    """G(-k) on the full grid: index i -> (-i) mod N on every axis (the
    single-process form of ``space_inversion_transposed_complex``,
    ``math_evaluator.py:806-977``, with ``_get_flip_partner`` :618-629)."""
    n = G.shape[0]
    inv = (-np.arange(n)) % n
    return G[np.ix_(inv, inv, inv)]


def high_order_strategy(poles, high_order_mode, correlation_mode):
    """``get_high_order_strategy`` (``clustering_estimator.py:243-331``)."""
    supported = ([0, 2, 4], [0, 2, 4, 6], [0, 2, 4, 6, 8])
    if high_order_mode == "default" or correlation_mode != "auto" or poles not in supported:
        return None
    high = {p for p in poles if p in (4, 6, 8)}
    fast, exact, internal = set(), set(), set()
    if high_order_mode == "fast":
        fast = {(0, 2, 4): {4}, (0, 2, 4, 6): {6}, (0, 2, 4, 6, 8): {6, 8}}.get(tuple(poles), set())
        exact = {(0, 2, 4): set(), (0, 2, 4, 6): {4}, (0, 2, 4, 6, 8): {4}}.get(tuple(poles), set())
    elif high_order_mode == "compare":
        fast, exact = set(high), set(high)
    if any(p in fast for p in (6, 8)):
        internal.add(4)
    return {"mode": high_order_mode, "compute_fast_poles": fast, "compute_exact_poles": exact | internal,
            "visible_exact_poles": exact, "need_f2_cache": any(p in fast for p in (4, 6, 8)),
            "need_f4_cache": any(p in fast for p in (6, 8))}


def calculate_power_spectrum_survey(stat_attrs, rfield_a, rfield_b, correlation_mode):
    """``clustering_estimator.py:205-563``."""
    poles = stat_attrs["poles"]
    n = int(np.atleast_1d(stat_attrs["nmesh"])[0])
    L = float(np.atleast_1d(stat_attrs["boxsize"])[0])
    boxcenter = np.array(stat_attrs["boxcenter"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler, interlaced = stat_attrs["sampler"], stat_attrs["interlaced"]
    N0 = stat_attrs["N0"]
    I_norm = stat_attrs["I_rand"] if stat_attrs["normalization_scheme"] == "particle" else stat_attrs["I_mesh"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    strat = high_order_strategy(poles, stat_attrs.get("high_order_mode", "default"), correlation_mode)
    comp = A.get_compensation(interlaced, sampler)
    w = C.circular_axes(n)
    V = L ** 3
    ca = comp(w, C.r2c(rfield_a)) * V
    cb = ca if correlation_mode == "auto" else comp(w, C.r2c(rfield_b)) * V
    kgrid, knorm = A.get_kgrid(C.k_axes(n, L))
    xgrid = A.get_xgrid(C.x_axes(n, L), boxcenter, np.array([L] * 3), np.array([n] * 3))
    k_num, ksum = A.get_kbin_count(k_bins, k_edge, knorm)
    res = {"k_eff": ksum / k_num, "k_num": k_num, "I_norm": I_norm}
    F2, F4 = None, None
    for pole in poles:
        if pole == 0:
            if correlation_mode == "auto":
                P = np.real(ca) ** 2 + np.imag(ca) ** 2
            else:
                P = ca * np.conj(cb)
        else:
            store = True
            if pole in (4, 6, 8) and strat is not None:
                store = pole in strat["compute_exact_poles"]
            cache2 = pole == 2 and strat is not None and strat["need_f2_cache"]
            cache4 = pole == 4 and strat is not None and strat["need_f4_cache"]
            if not store and not cache2 and not cache4:
                continue
            G = get_G_ell(rfield_b, pole, kgrid, xgrid, comp, w, L)
            if cache2:
                F2 = (G + np.conj(space_inversion(G))) * (1 / 2)
            if cache4:
                F4 = (G + np.conj(space_inversion(G))) * (1 / 2)
            if not store:
                continue
            P = ca * np.conj(G)
        s = A.radial_binning(P, k_bins, k_edge, knorm)
        r = s / k_num
        r = r * ((2 * pole + 1) / I_norm)
        if pole == 0:
            r = r - N0
        else:
            r = np.imag(r) if pole & 1 else np.real(r)
        res[f"P{pole}"] = r
    if strat is not None and strat["compute_fast_poles"]:
        fast = strat["compute_fast_poles"]
        if 4 in fast:
            P22 = A.radial_binning(F2 * np.conj(F2), k_bins, k_edge, knorm) / k_num
            P22 = P22 * (9 / I_norm) - N0 * 9 / 5
            res["P4b"] = np.real(35 / 18 * P22 - res["P2"] - 7 / 2 * res["P0"])
        if 6 in fast or 8 in fast:
            s42 = A.radial_binning(F4 * np.conj(F2), k_bins, k_edge, knorm)
            if 6 in fast:
                P42 = s42 / k_num * (13 / I_norm)
                res["P6b"] = np.real(11 / 5 * P42 - 52 / 63 * res["P4"] - 286 / 175 * res["P2"])
            if 8 in fast:
                P42 = s42 / k_num * (13 / I_norm)
                P44 = A.radial_binning(F4 * np.conj(F4), k_bins, k_edge, knorm) / k_num * (17 / I_norm)
                P44 = P44 - N0 * 17 / 9
                res["P8b"] = np.real(1287 / 490 * P44 - 374 / 245 * P42 - 3553 / 15435 * res["P4"]
                                     - 1326 / 8575 * res["P2"] - 2431 / 490 * res["P0"])
        if strat["mode"] == "fast":
            for pole in fast:
                res.pop(f"P{pole}", None)
    return res


# ---------------------------------------------------------------- bk meshes
def _prepare_tracer(cats, dkey, rkey, corr, angu_config, n, L, sampler, interlaced, boxcenter, alpha_scheme):
    """``get_mesh_bk_survey._prepare_tracer`` (``mesh_generator.py:589-681``)."""
    vpc = L ** 3 / n ** 3
    d, r = cats[dkey], cats[rkey]
    sums = compute_particle_sums_bk(d, r, corr, angu_config)
    alpha = _alpha(sums, alpha_scheme)
    I_33 = sums[4] * alpha
    S_LM = sums[5] - alpha ** 3 * sums[6]
    rd = _paint(d, d["WEIGHT"] * d["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
    rr = _paint(r, r["WEIGHT"] * r["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
    rd = rd * (1 / vpc)
    rr = rr * (alpha / vpc)
    return alpha, I_33, S_LM, rd - rr, rd, rr


def get_mesh_bk_survey(cats, correlation_mode, tracer_type, angu_config, nmesh, boxsize, sampler, interlaced,
                       normalization_scheme="particle", alpha_scheme="pypower", boxcenter=None):
    """``mesh_generator.py:541-1004`` with catalogs already read.

    Returns (mesh_attrs, rfield_a, rfield_b, rfield_c)."""
    n, L = int(nmesh), float(boxsize)
    vpc = L ** 3 / n ** 3
    if correlation_mode == "auto":
        assert tracer_type == "aaa", "For auto-correlation, tracer_type must be 'aaa'."
    attrs = {"normalization_scheme": normalization_scheme}
    if boxcenter is None:
        boxcenter = survey_boxcenter(cats["randoms_a"])
    attrs["boxcenter"] = boxcenter
    fields = {}
    labels = ["a"] if correlation_mode == "auto" else (["a", "b", "c"] if tracer_type == "abc" else ["a", "b"])
    for lab in labels:
        alpha, I_33, S_LM, F, rd, rr = _prepare_tracer(cats, f"data_{lab}", f"randoms_{lab}", correlation_mode,
                                                       angu_config, n, L, sampler, interlaced, boxcenter,
                                                       alpha_scheme)
        attrs[f"alpha_{lab}"] = alpha
        attrs[f"I_33_{lab}"] = I_33
        attrs[f"S_LM_{lab}"] = S_LM
        fields[lab] = (F, rd, rr)
    if normalization_scheme == "particle":
        assert correlation_mode == "auto", "Particle-normalization scheme is only applicable for auto-correlation."
        attrs["I_norm"] = attrs["I_33_a"]
    elif normalization_scheme == "mesh":
        assert correlation_mode == "auto", "Mesh-normalization scheme is only applicable for auto-correlation."
        attrs["I_norm"] = vpc * np.sum(fields["a"][2] ** 3)
    else:
        if correlation_mode == "auto":
            d, r = cats["data_a"], cats["randoms_a"]
            odd = np.arange(r["Position"].shape[0]) % 2 == 1
            r_odd, r_even = r[odd], r[~odd]
            if alpha_scheme == "pypower":
                sd = np.sum(d["WEIGHT"] * d["WEIGHT_FKP"])
                so = np.sum(r_odd["WEIGHT"] * r_odd["WEIGHT_FKP"])
                se = np.sum(r_even["WEIGHT"] * r_even["WEIGHT_FKP"])
            else:
                sd, so, se = np.sum(d["WEIGHT"]), np.sum(r_odd["WEIGHT"]), np.sum(r_even["WEIGHT"])
            a_odd, a_even = sd / so, sd / se
            ro = _paint(r_odd, r_odd["WEIGHT"] * r_odd["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
            re = _paint(r_even, r_even["WEIGHT"] * r_even["WEIGHT_FKP"], n, L, sampler, interlaced, boxcenter)
            ro = ro * (a_odd / vpc)
            re = re * (a_even / vpc)
            attrs["I_norm"] = vpc * np.sum(fields["a"][1] * ro * re)
        elif tracer_type == "abc":
            attrs["I_norm"] = vpc * np.sum(fields["a"][2] * fields["b"][2] * fields["c"][2])
        elif tracer_type == "aab":
            attrs["I_norm"] = vpc * np.sum(fields["a"][1] * fields["a"][2] * fields["b"][2])
        else:
            attrs["I_norm"] = vpc * np.sum(fields["a"][2] * fields["b"][2] * fields["b"][1])
    fa = fields["a"][0]
    fb = fields["b"][0] if "b" in fields else None
    fc = fields["c"][0] if "c" in fields else None
    return attrs, fa, fb, fc


def get_N_field(cats, tracer_flag, alpha, nmesh, boxsize, sampler, interlaced, boxcenter):
    """``mesh_generator.py:1007-1128``: N = D_{w^2}/dV + alpha^2 R_{w^2}/dV."""
    n, L = int(nmesh), float(boxsize)
    vpc = L ** 3 / n ** 3
    d, r = cats[f"data_{tracer_flag}"], cats[f"randoms_{tracer_flag}"]
    rd = _paint(d, (d["WEIGHT"] * d["WEIGHT_FKP"]) ** 2, n, L, sampler, interlaced, boxcenter)
    rr = _paint(r, (r["WEIGHT"] * r["WEIGHT_FKP"]) ** 2, n, L, sampler, interlaced, boxcenter)
    rd = rd / vpc
    rr = rr / vpc
    return rd + alpha ** 2 * rr


# ---------------------------------------------------------------- bk estimator
def _fill_and_accumulate(total, sub, cfg, use_sym, sym_eq, sym_cj, add_swap, diagonal):
    """Symmetric fill, swap reuse and the (0,0,0)-once / 2 Re rule
    (``clustering_estimator.py:1858-1913``, ``:2024-2052``)."""
    if not diagonal and use_sym:
        kb = sub.shape[0]
        for i in range(kb):
            for j in range(i + 1, kb):
                sub[j, i] = sub[i, j] if sym_eq else np.conj(sub[i, j])
    add = sub
    if add_swap:
        add = 2.0 * sub if diagonal else sub + sub.T
    if cfg == (0, 0, 0):
        return total + add
    return total + 2 * np.real(add)


def calculate_bk_sugi_survey(rfield_a, rfield_b, rfield_c, correlation_mode, stat_attrs, cats,
                             return_signal_only=False):
    """``clustering_estimator.py:1558-3101`` (MPI size 1).

    With ``return_signal_only`` the signal-stage ``total_res`` before
    normalisation is returned."""
    dvm = stat_attrs.get("data_vector_mode", "diagonal")
    sn_mode = stat_attrs.get("shotnoise_mode", "ana")
    block_size = stat_attrs.get("block_size", 1 if dvm == "diagonal" else "full")
    ell_1, ell_2, L = stat_attrs["angu_config"]
    n = int(np.atleast_1d(stat_attrs["nmesh"])[0])
    Lbox = float(np.atleast_1d(stat_attrs["boxsize"])[0])
    boxcenter = np.array(stat_attrs["boxcenter"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler = stat_attrs["sampler"]
    interlaced = stat_attrs["interlaced"]
    tracer_type = stat_attrs["tracer_type"]
    I_norm = stat_attrs["I_norm"]
    V = Lbox ** 3
    vpc = V / n ** 3
    N_ells = (2 * ell_1 + 1) * (2 * ell_2 + 1) * (2 * L + 1)
    H_ells = A.wigner_3j(ell_1, ell_2, L, 0, 0, 0)
    mcfg, tjv = A.get_magnetic_configs_survey(ell_1, ell_2, L)
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    w = C.circular_axes(n)

    def comp(v):
        return A.compensate_sinc(w, v, sampler)

    def comp_shot(v):
        return A.compensate_bk_noise(w, v, sampler)

    ca = comp(C.r2c(rfield_a)) * V
    cb = ca if correlation_mode == "auto" else comp(C.r2c(rfield_b)) * V
    kgrid, knorm = A.get_kgrid(C.k_axes(n, Lbox))
    xgrid = A.get_xgrid(C.x_axes(n, Lbox), boxcenter, np.array([Lbox] * 3), np.array([n] * 3))
    k_num, ksum = A.get_kbin_count(k_bins, k_edge, knorm)
    k_eff = ksum / k_num
    if correlation_mode == "auto":
        rf3 = rfield_a
    else:
        rf3 = rfield_c if tracer_type == "abc" else rfield_b

    def G_LM(M):
        y = A.get_Ylm(L, M, Racah_normalized=True)
        return C.c2r(comp(C.r2c(rf3 * y(*xgrid))))

    def mask(i):
        return np.logical_and(knorm >= k_edge[i], knorm < k_edge[i + 1])

    cfg_idx = {c: i for i, c in enumerate(mcfg)}
    processed = set()
    same_tracer_12 = correlation_mode == "auto" or tracer_type == "aab"
    same_ell = ell_1 == ell_2
    cur_M, G = None, None
    if dvm == "full":
        if block_size == "full":
            block_size = k_bins
        elif not (isinstance(block_size, int) and 1 <= block_size <= k_bins):
            raise ValueError("For full mode, block_size must be 1, 'full', or an integer in [1, k_bins].")
        total = np.zeros((k_bins, k_bins)).astype("complex128")
        for ss, (m1, m2, M) in enumerate(mcfg):
            if (m1, m2, M) in processed:
                continue
            sub = np.zeros((k_bins, k_bins)).astype("complex128")
            if M != cur_M:
                G, cur_M = G_LM(M), M
            w1 = ca * A.get_Ylm(ell_1, m1, True)(*kgrid)
            if tracer_type in ["aaa", "aab"]:
                w2 = w1 if (ell_1 == ell_2 and m1 == m2) else ca * A.get_Ylm(ell_2, m2, True)(*kgrid)
            else:
                w2 = cb * A.get_Ylm(ell_2, m2, True)(*kgrid)
            sym_eq = same_tracer_12 and same_ell and m1 == m2
            sym_cj = same_tracer_12 and same_ell and m1 == -m2
            use_sym = sym_eq or sym_cj
            F1 = [C.c2r(w1 * mask(i)) for i in range(k_bins)]
            for i in range(k_bins):
                for j in range(i if use_sym else 0, k_bins):
                    if sym_eq:
                        F2 = F1[j]
                    elif sym_cj:
                        F2 = (-1) ** (m1 + ell_1) * np.conj(F1[j])
                    else:
                        F2 = C.c2r(w2 * mask(j))
                    sub[i, j] = tjv[ss] * np.sum(G * F1[i] * F2)
            swap = (m2, m1, M)
            has_swap = same_tracer_12 and same_ell and swap in cfg_idx and swap != (m1, m2, M)
            add_swap = has_swap and swap not in processed
            total = _fill_and_accumulate(total, sub, (m1, m2, M), use_sym, sym_eq, sym_cj, add_swap, False)
            if add_swap:
                processed.add(swap)
            processed.add((m1, m2, M))
    elif dvm == "diagonal":
        total = np.zeros(k_bins).astype("complex128")
        reuse_fixed = ell_2 == 0 and all(c[1] == 0 for c in mcfg)
        shared2 = None
        if reuse_fixed:
            shared2 = (ca if tracer_type in ["aaa", "aab"] else cb) * A.get_Ylm(ell_2, 0, True)(*kgrid)
        for ss, (m1, m2, M) in enumerate(mcfg):
            if (m1, m2, M) in processed:
                continue
            sub = np.zeros(k_bins).astype("complex128")
            if M != cur_M:
                G, cur_M = G_LM(M), M
            w1 = ca * A.get_Ylm(ell_1, m1, True)(*kgrid)
            sym_eq = same_tracer_12 and same_ell and m1 == m2
            sym_cj = same_tracer_12 and same_ell and m1 == -m2
            w2 = None
            if not (sym_eq or sym_cj):
                if reuse_fixed:
                    w2 = shared2
                else:
                    w2 = (ca if tracer_type in ["aaa", "aab"] else cb) * A.get_Ylm(ell_2, m2, True)(*kgrid)
            for i in range(k_bins):
                F1 = C.c2r(w1 * mask(i))
                if sym_eq:
                    F2 = F1
                elif sym_cj:
                    F2 = (-1) ** (m1 + ell_1) * np.conj(F1)
                else:
                    F2 = C.c2r(w2 * mask(i))
                sub[i] = tjv[ss] * np.sum(G * F1 * F2)
            swap = (m2, m1, M)
            has_swap = same_tracer_12 and same_ell and swap in cfg_idx and swap != (m1, m2, M)
            add_swap = has_swap and swap not in processed
            total = _fill_and_accumulate(total, sub, (m1, m2, M), False, sym_eq, sym_cj, add_swap, True)
            if add_swap:
                processed.add(swap)
            processed.add((m1, m2, M))
    else:
        raise ValueError("data_vector_mode must be either 'diagonal' or 'full'.")
    if return_signal_only:
        return total

    # ------------------------------------------------------------ shot noise
    rf_shot, N_shot, S_LM_shot = None, None, None
    if tracer_type != "abc":
        nm = stat_attrs["nmesh"]
        flag = "b" if tracer_type == "abb" else "a"
        N_shot = get_N_field(cats, flag, stat_attrs[f"alpha_{flag}"], n, Lbox, sampler, interlaced, boxcenter)
        rf_shot = rfield_b if tracer_type == "aab" else rfield_a
        S_LM_shot = stat_attrs.get(f"S_LM_{flag}", None)
        del nm
    k_center = 0.5 * (k_edge[1:] + k_edge[:-1])
    k3_edge = np.linspace(2 * k_min, 2 * k_max, k_bins * 2 + 1)[:-1]
    k3_center = 0.5 * (k3_edge[1:] + k3_edge[:-1])
    k3_bins = len(k3_center)
    k_num_k3, ksum3 = A.get_kbin_count(k3_bins, k3_edge, knorm)
    with np.errstate(invalid="ignore", divide="ignore"):
        k3_eff = ksum3 / k_num_k3
    use_sbar = correlation_mode == "auto" and tracer_type == "aaa"

    def s_lm_array(S):
        if S is None:
            return np.zeros(L + 1, dtype="complex128")
        if np.isscalar(S):
            a = np.zeros(L + 1, dtype="complex128")
            a[0] = S
            return a
        return np.asarray(S, dtype="complex128")

    def S1_like():
        """``clustering_estimator.py:2231-2320``."""
        out = np.zeros(k_bins, dtype="complex128")
        if not (ell_1 == L and ell_2 == 0) or rf_shot is None or N_shot is None:
            return out
        cd = comp(C.r2c(rf_shot)) * V
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None
        acc = np.zeros_like(cd, dtype="complex128")
        for M in range(0, L + 1):
            mw = 0.5 if M == 0 else 1.0
            y = A.get_Ylm(L, M, Racah_normalized=True)
            cn = comp(C.r2c(N_shot * y(*xgrid))) * V
            term = cd * np.conj(cn)
            if use_sbar:
                term = term - comp_shot(np.full(cd.shape, S_arr[M], dtype=np.complex128))
            term = term * y(*kgrid)
            acc = acc + mw * term
        s = 2.0 * np.real(A.radial_binning(acc, k_bins, k_edge, knorm))
        ok = k_num > 0
        out[ok] = (2 * L + 1) * s[ok] / k_num[ok]
        return out

    def Q_L_like():
        """``clustering_estimator.py:2322-2401``."""
        if rf_shot is None or N_shot is None:
            return np.zeros(k3_bins, dtype="complex128")
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None
        cn0 = comp(C.r2c(N_shot)) * V
        acc = np.zeros_like(cn0, dtype="complex128")
        for M in range(0, L + 1):
            mw = 0.5 if M == 0 else 1.0
            y = A.get_Ylm(L, M, Racah_normalized=True)
            cdl = comp(C.r2c(rf_shot * np.conj(y(*xgrid)))) * V
            term = cdl * np.conj(cn0)
            if use_sbar:
                term = term - comp_shot(np.full(cdl.shape, S_arr[M], dtype=np.complex128))
            term = term * y(*kgrid)
            acc = acc + mw * term
        return (2.0 * np.real(A.radial_binning(acc, k3_bins, k3_edge, knorm))).astype("complex128")

    def S3_ana(Q, mode):
        """``clustering_estimator.py:2403-2491``."""
        wkn = np.zeros(k3_bins, dtype="float64")
        ok = k_num_k3 > 0
        wkn[ok] = k_num_k3[ok] / k3_center[ok]
        if mode == "diagonal":
            s3 = np.zeros(k_bins, dtype="complex128")
            for i in range(k_bins):
                q = A.get_q_ells_eff(i, i, k_eff, ell_1, ell_2, L, k3_bins, k3_eff)
                numer = np.sum(q * Q / k3_center)
                denom = np.sum(wkn[: 2 * i + 1])
                if denom > 0:
                    s3[i] = numer / denom
        else:
            s3 = np.zeros((k_bins, k_bins), dtype="complex128")
            for i in range(k_bins):
                for j in range(k_bins):
                    q = A.get_q_ells_eff(i, j, k_eff, ell_1, ell_2, L, k3_bins, k3_eff)
                    valid = A.get_valid_k3_bins(k_center[i], k_center[j], k_min, k_max, k_bins)
                    numer = np.sum(q * Q / k3_center)
                    denom = np.sum(wkn * valid)
                    if denom > 0:
                        s3[i, j] = numer / denom
        return s3 * (H_ells * N_ells)

    def S3_fft(mode):
        """``clustering_estimator.py:2510-2781``."""
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None

        def s_lm_value(m_idx):
            if S_arr is None:
                return 0.0 + 0.0j
            if m_idx >= 0:
                return S_arr[m_idx]
            return ((-1) ** m_idx) * np.conj(S_arr[-m_idx])

        def shell(ell, m, b):
            return C.c2r(np.conj(A.get_Ylm(ell, m, True)(*kgrid)) * mask(b))

        cn0 = comp(C.r2c(N_shot)) * V

        def source(M):
            y = A.get_Ylm(L, M, Racah_normalized=True)
            cdl = comp(C.r2c(rf_shot * np.conj(y(*xgrid)))) * V
            src = cdl * np.conj(cn0)
            if use_sbar:
                src = src - comp_shot(np.full(cdl.shape, s_lm_value(M), dtype=np.complex128))
            return C.c2r(src)

        cur, src = None, None
        done = set()
        s3 = np.zeros(k_bins if mode == "diagonal" else (k_bins, k_bins), dtype="complex128")
        for ss, (m1, m2, M) in enumerate(mcfg):
            if (m1, m2, M) in done:
                continue
            if M != cur:
                src, cur = source(M), M
            sym_eq = same_ell and m1 == m2
            sym_cj = same_ell and m1 == -m2
            if mode == "diagonal":
                sub = np.zeros(k_bins, dtype="complex128")
                for i in range(k_bins):
                    f1 = shell(ell_1, m1, i)
                    if sym_eq:
                        f2 = f1
                    elif sym_cj:
                        f2 = ((-1) ** (m1 + ell_1)) * np.conj(f1)
                    else:
                        f2 = shell(ell_2, m2, i)
                    if k_num[i] > 0:
                        sub[i] = tjv[ss] * np.sum(f1 * f2 * src) / (k_num[i] ** 2)
                use_sym = False
            else:
                sub = np.zeros((k_bins, k_bins), dtype="complex128")
                use_sym = sym_eq or sym_cj
                f1s = [shell(ell_1, m1, i) for i in range(k_bins)]
                for i in range(k_bins):
                    for j in range(i if use_sym else 0, k_bins):
                        if sym_eq:
                            f2 = f1s[j]
                        elif sym_cj:
                            f2 = ((-1) ** (m1 + ell_1)) * np.conj(f1s[j])
                        else:
                            f2 = shell(ell_2, m2, j)
                        if k_num[i] > 0 and k_num[j] > 0:
                            sub[i, j] = tjv[ss] * np.sum(f1s[i] * f2 * src) / (k_num[i] * k_num[j])
            swap = (m2, m1, M)
            has_swap = same_ell and swap in cfg_idx and swap != (m1, m2, M)
            add_swap = has_swap and swap not in done
            s3 = _fill_and_accumulate(s3, sub, (m1, m2, M), use_sym, sym_eq, sym_cj, add_swap, mode == "diagonal")
            if add_swap:
                done.add(swap)
            done.add((m1, m2, M))
        return s3 * (H_ells * N_ells) / float(n) ** 3

    def S3(mode, which):
        if tracer_type == "abb":
            return np.zeros((k_bins, k_bins) if mode == "full" else k_bins, dtype="complex128")
        if which == "ana":
            return S3_ana(Q_L_like(), mode)
        return S3_fft(mode)

    shape = (k_bins,) if dvm == "diagonal" else (k_bins, k_bins)
    SN0 = 0.0 + 0.0j
    SN1 = np.zeros(shape, dtype="complex128")
    SN2 = np.zeros(shape, dtype="complex128")
    SN3 = np.zeros(shape, dtype="complex128")
    SN3_fft = None
    if tracer_type != "abc":
        if tracer_type == "aaa" and (ell_1, ell_2, L) == (0, 0, 0):
            S0 = stat_attrs.get("S_LM_a", 0.0)
            SN0 = S0[0] if isinstance(S0, np.ndarray) else S0
        s1 = S1_like() if (tracer_type in ["aaa", "abb"] and ell_1 == L and ell_2 == 0) \
            else np.zeros(k_bins, dtype="complex128")
        s2 = s1.copy() if (tracer_type == "aaa" and (ell_1, ell_2, L) == (0, 0, 0)) \
            else np.zeros(k_bins, dtype="complex128")
        if dvm == "diagonal":
            SN1, SN2 = s1, s2
        else:
            for i in range(k_bins):
                SN1[i, :] = s1[i]
            for i in range(k_bins):
                SN2[:, i] = s2[i]
        if tracer_type != "abb":
            if sn_mode == "both":
                SN3 = S3(dvm, "ana")
                SN3_fft = S3(dvm, "fft")
            else:
                SN3 = S3(dvm, sn_mode)
        elif sn_mode == "both":
            SN3_fft = np.zeros(shape, dtype="complex128")
    total = total * (N_ells * H_ells / I_norm)
    if dvm == "full":
        total = total * (1.0 / np.outer(k_num, k_num))
    else:
        total = total * (1.0 / (k_num ** 2))
    total = total * vpc
    tsn = (SN0 + SN1 + SN2 + SN3) / I_norm
    sn_terms = {"SN0": SN0 / I_norm, "SN1": SN1 / I_norm, "SN2": SN2 / I_norm, "SN3": SN3 / I_norm}
    if sn_mode == "both":
        sn_terms["SN3_FFT"] = SN3_fft / I_norm
    return {"B_sugi": total - tsn, "Bk_raw": total, "SN_terms": sn_terms, "Shot_noise": tsn,
            "I_norm": I_norm, "k_eff": k_eff, "nmodes": k_num}
