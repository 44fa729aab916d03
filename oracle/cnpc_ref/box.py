"""Box-geometry estimators restated on full complex128 grids (oracle; test
infrastructure only).

Follows ``get_mesh_box`` (``mesh_generator.py:15-169``),
``calculate_power_spectrum_box`` (``clustering_estimator.py:566-746``) and
``calculate_bk_sugi_box`` (``clustering_estimator.py:749-1555``) with MPI size 1.
Catalogs are passed as arrays (the file readers are outside the hot path).
"""
import numpy as np

from . import angular as A
from . import conventions as C


def get_mesh_box(cats, nmesh, boxsize, sampler, interlaced, correlation_mode="auto", tracer_type="aaa"):
    """``mesh_generator.py:15-169``: N_gal and NZ = sum(W)/V per tracer
    (``:56-60``), then one painted (interlaced) field per tracer.
    ``cats`` maps 'a'/'b'/'c' to (Position (Np,3) f8, WEIGHT (Np,) f8)."""
    V = float(np.prod(boxsize))
    attrs = {}
    fields = {}
    need = ["a"]
    if correlation_mode == "cross":
        need.append("b")
    if tracer_type == "abc":
        need.append("c")
    for t in need:
        pos, w = cats[t]
        attrs[f"N_gal_{t}"] = len(w)
        attrs[f"NZ_{t}"] = np.sum(w) / V
        fields[t] = C.pm_painter(pos, w, nmesh[0], boxsize[0], sampler, interlaced)
    attrs.update({"interlaced": interlaced, "sampler": sampler})
    return attrs, fields.get("a"), fields.get("b"), fields.get("c")


def calculate_power_spectrum_box(rfield_a, rfield_b, correlation_mode, stat_attrs):
    """``clustering_estimator.py:566-746``."""
    poles = stat_attrs["poles"]
    boxsize, nmesh = np.array(stat_attrs["boxsize"]), np.array(stat_attrs["nmesh"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler, interlaced = stat_attrs["sampler"], stat_attrs["interlaced"]
    rsd = np.array(stat_attrs["rsd"])
    NZ_a = stat_attrs["NZ_a"]
    NZ_b = stat_attrs.get("NZ_b")
    P_shot = 1.0 / NZ_a if correlation_mode == "auto" else 0.0
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    comp = A.get_compensation(interlaced, sampler)
    w = C.circular_axes(nmesh[0])
    ca = comp(w, C.r2c(rfield_a))
    cb = comp(w, C.r2c(rfield_b)) if correlation_mode == "cross" else ca
    kax = C.k_axes(nmesh[0], boxsize[0])
    kgrid, knorm = A.get_kgrid(kax)
    if knorm[0, 0, 0] == 0.0:
        ca = ca.copy()
        ca[0, 0, 0] = 0.0
        if correlation_mode != "cross":
            cb = ca
    P_field = np.real(ca) * np.real(cb) + np.imag(ca) * np.imag(cb)
    count, ksum = A.get_kbin_count(k_bins, k_edge, knorm)
    k_num = count
    results = {"k_eff": ksum / k_num, "k_num": k_num}
    for pole in poles:
        if pole & 1:
            res = np.zeros(k_bins).astype("float128")
        else:
            if pole == 0:
                s = A.radial_binning(P_field, k_bins, k_edge, knorm)
            else:
                Lf = A.get_legendre(pole, rsd[0], rsd[1], rsd[2])
                s = A.radial_binning(P_field * Lf(kgrid[0], kgrid[1], kgrid[2]), k_bins, k_edge, knorm)
            res = s / k_num
            if correlation_mode == "auto":
                res *= boxsize.prod() ** (-1) * (nmesh.prod() ** 2) / NZ_a ** 2
            else:
                res *= boxsize.prod() ** (-1) * nmesh.prod() ** 2 / (NZ_a * NZ_b)
            if pole == 0:
                res -= P_shot
            else:
                res *= 2 * pole + 1
        results.update({f"P{pole}": res, "P_shot": P_shot})
    return results


def _bk_prepare(rfield, NZ, vol_per_cell, sampler, nmesh):
    """``clustering_estimator.py:808-816``: delta = n/dV - NZ, r2c, sinc**p."""
    d = rfield / vol_per_cell - NZ
    return A.compensate_sinc(C.circular_axes(nmesh), C.r2c(d), sampler)


def calculate_bk_sugi_box(rfield_a, rfield_b, rfield_c, correlation_mode, stat_attrs, return_signal_only=False):
    """``clustering_estimator.py:749-1555`` (MPI size 1).

    With ``return_signal_only`` the k_bins x k_bins (or k_bins) ``total_res``
    before normalisation (``:991-995`` / ``:1057-1061``) is returned, which is
    the signal-stage output the GPU contraction is compared with."""
    dvm = stat_attrs.get("data_vector_mode", "diagonal")
    sn_mode = stat_attrs.get("shotnoise_mode", "ana")
    all_eff = stat_attrs.get("shotnoise_ana_all_eff", True)
    block_size = stat_attrs.get("block_size", 1 if dvm == "diagonal" else "full")
    ell_1, ell_2, L = stat_attrs["angu_config"]
    boxsize = np.array(stat_attrs["boxsize"])
    nmesh = np.array(stat_attrs["nmesh"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler = stat_attrs["sampler"]
    tracer_type = stat_attrs["tracer_type"]
    vpc = boxsize.prod() / nmesh.prod()
    N = int(nmesh[0])

    N_ells = (2 * ell_1 + 1) * (2 * ell_2 + 1) * (2 * L + 1)
    H_ells = A.wigner_3j(ell_1, ell_2, L, 0, 0, 0)
    mcfg, tjv = A.get_magnetic_configs_box(ell_1, ell_2, L)

    NZ_a, N_gal_a = stat_attrs["NZ_a"], stat_attrs["N_gal_a"]
    if correlation_mode == "cross":
        NZ_b, N_gal_b = stat_attrs["NZ_b"], stat_attrs["N_gal_b"]
        if tracer_type == "abc":
            NZ_c, N_gal_c = stat_attrs["NZ_c"], stat_attrs["N_gal_c"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)

    ca = _bk_prepare(rfield_a, NZ_a, vpc, sampler, N)
    cb = _bk_prepare(rfield_b, NZ_b, vpc, sampler, N) if correlation_mode == "cross" else None
    cc = _bk_prepare(rfield_c, NZ_c, vpc, sampler, N) if (correlation_mode == "cross" and tracer_type == "abc") else None

    kgrid, knorm = A.get_kgrid(C.k_axes(N, boxsize[0]))
    if correlation_mode == "auto":
        G_00 = C.c2r(ca)
    else:
        G_00 = C.c2r(cc if tracer_type == "abc" else cb)
    G_00 = np.real(G_00)

    k_num, ksum = A.get_kbin_count(k_bins, k_edge, knorm)
    k_eff = ksum / k_num

    def mask(i):
        return np.logical_and(knorm >= k_edge[i], knorm < k_edge[i + 1])

    if dvm == "full":
        if block_size == "full":
            block_size = k_bins
        elif not (isinstance(block_size, int) and 1 <= block_size <= k_bins):
            raise ValueError("For full mode, block_size must be 1, 'full', or an integer in [1, k_bins].")
        total_res = np.zeros((k_bins, k_bins)).astype("complex128")
        for ss, (m1, m2, M) in enumerate(mcfg):
            sub = np.zeros((k_bins, k_bins)).astype("complex128")
            y1 = A.get_Ylm(ell_1, m1, True)(*kgrid)
            y2 = A.get_Ylm(ell_2, m2, True)(*kgrid)
            w1 = ca * y1
            w2 = (ca if tracer_type in ["aaa", "aab"] else cb) * y2
            same_tracer = correlation_mode == "auto" or tracer_type == "aab"
            same_ell = ell_1 == ell_2
            sym_eq = same_tracer and same_ell and (m1 == m2)
            sym_cj = same_tracer and same_ell and (m1 == -m2)
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
                    sub[i, j] = tjv[ss] * np.sum(G_00 * F1[i] * F2)
            if use_sym:
                for i in range(k_bins):
                    for j in range(i + 1, k_bins):
                        sub[j, i] = sub[i, j] if sym_eq else np.conj(sub[i, j])
            total_res += sub if (m1, m2, M) == (0, 0, 0) else 2 * np.real(sub)
    elif dvm == "diagonal":
        total_res = np.zeros(k_bins).astype("complex128")
        for ss, (m1, m2, M) in enumerate(mcfg):
            sub = np.zeros(k_bins).astype("complex128")
            y1 = A.get_Ylm(ell_1, m1, True)(*kgrid)
            w1 = ca * y1
            if tracer_type in ["aaa", "aab"]:
                if ell_1 != ell_2 or m1 != m2:
                    w2 = ca * A.get_Ylm(ell_2, m2, True)(*kgrid)
            else:
                w2 = cb * A.get_Ylm(ell_2, m2, True)(*kgrid)
            for i in range(k_bins):
                mk = mask(i)
                F1 = C.c2r(w1 * mk)
                if tracer_type in ["aaa", "aab"] and ell_1 == ell_2:
                    F2 = F1 if m1 == 0 else (-1) ** (m1 + ell_1) * np.conj(F1)
                else:
                    F2 = C.c2r(w2 * mk)
                sub[i] = tjv[ss] * np.sum(G_00 * F1 * F2)
            total_res += sub if (m1, m2, M) == (0, 0, 0) else 2 * np.real(sub)
    else:
        raise ValueError("data_vector_mode must be either 'diagonal' or 'full'")
    if return_signal_only:
        return total_res

    # k3 bins (``:1063-1074``)
    k_center = 0.5 * (k_edge[1:] + k_edge[:-1])
    k3_edge = np.linspace(2 * k_min, 2 * k_max, k_bins * 2 + 1)[:-1]
    k3_center = 0.5 * (k3_edge[1:] + k3_edge[:-1])
    k3_bins = len(k3_center)
    k_num_k3, ksum3 = A.get_kbin_count(k3_bins, k3_edge, knorm)
    with np.errstate(invalid="ignore", divide="ignore"):
        k3_eff = ksum3 / k_num_k3
    w_circ = C.circular_axes(N)

    def s3_ana(total_sum):
        """``:1081-1180``."""
        if tracer_type == "abb":
            return 0.0
        weighted_k_num = k_num_k3 / k3_center
        if dvm == "diagonal":
            s3 = np.zeros(k_bins, dtype="complex128")
            for i in range(k_bins):
                if all_eff:
                    q = A.get_q_ells_eff(i, i, k_eff, ell_1, ell_2, L, k3_bins, k3_eff)
                else:
                    q = A.get_q_ells(i, i, k_center, k_min, k_max, k_bins, ell_1, ell_2, L, k3_bins)
                numer = np.sum(q * total_sum / k3_center)
                denom = np.sum(weighted_k_num[: 2 * i + 1])
                if denom > 0:
                    s3[i] = numer / denom
        else:
            s3 = np.zeros((k_bins, k_bins), dtype="complex128")
            for i in range(k_bins):
                for j in range(k_bins):
                    if all_eff:
                        q = A.get_q_ells_eff(i, j, k_eff, ell_1, ell_2, L, k3_bins, k3_eff)
                    else:
                        q = A.get_q_ells(i, j, k_center, k_min, k_max, k_bins, ell_1, ell_2, L, k3_bins)
                    valid = A.get_valid_k3_bins(k_center[i], k_center[j], k_min, k_max, k_bins)
                    numer = np.sum(q * total_sum / k3_center)
                    denom = np.sum(weighted_k_num * valid)
                    if denom > 0:
                        s3[i, j] = numer / denom
        return s3 * (H_ells * N_ells)

    def s3_fft():
        """``:1182-1384``."""
        if tracer_type == "abb":
            return np.zeros((k_bins, k_bins) if dvm == "full" else k_bins, dtype="complex128")
        src = ca * np.conj(ca if correlation_mode == "auto" else cb) * boxsize.prod() ** 2
        if correlation_mode == "auto":
            src = src - A.compensate_bk_noise(w_circ, np.full(src.shape, SN0, dtype=np.complex128), sampler)
        src = C.c2r(src)

        def shell(ell, m, b):
            return C.c2r(np.conj(A.get_Ylm(ell, m, True)(*kgrid)) * mask(b))

        same_ell = ell_1 == ell_2
        if dvm == "diagonal":
            s3 = np.zeros(k_bins, dtype="complex128")
            for ss, (m1, m2, M) in enumerate(mcfg):
                sub = np.zeros(k_bins, dtype="complex128")
                for i in range(k_bins):
                    f1 = shell(ell_1, m1, i)
                    if same_ell and m1 == m2:
                        f2 = f1
                    elif same_ell and m1 == -m2:
                        f2 = (-1) ** (m1 + ell_1) * np.conj(f1)
                    else:
                        f2 = shell(ell_2, m2, i)
                    if k_num[i] > 0:
                        sub[i] = tjv[ss] * np.sum(src * f1 * f2) / k_num[i] ** 2
                s3 += sub if (m1, m2, M) == (0, 0, 0) else 2.0 * np.real(sub)
        else:
            s3 = np.zeros((k_bins, k_bins), dtype="complex128")
            for ss, (m1, m2, M) in enumerate(mcfg):
                sub = np.zeros((k_bins, k_bins), dtype="complex128")
                sym_eq = same_ell and m1 == m2
                sym_cj = same_ell and m1 == -m2
                use_sym = sym_eq or sym_cj
                f1s = [shell(ell_1, m1, i) for i in range(k_bins)]
                for i in range(k_bins):
                    for j in range(i if use_sym else 0, k_bins):
                        if sym_eq:
                            f2 = f1s[j]
                        elif sym_cj:
                            f2 = (-1) ** (m1 + ell_1) * np.conj(f1s[j])
                        else:
                            f2 = shell(ell_2, m2, j)
                        if k_num[i] > 0 and k_num[j] > 0:
                            sub[i, j] = tjv[ss] * np.sum(src * f1s[i] * f2) / (k_num[i] * k_num[j])
                if use_sym:
                    for i in range(k_bins):
                        for j in range(i + 1, k_bins):
                            sub[j, i] = sub[i, j] if sym_eq else np.conj(sub[i, j])
                s3 += sub if (m1, m2, M) == (0, 0, 0) else 2.0 * np.real(sub)
        return s3 * (H_ells * N_ells) / nmesh.prod()

    SN3_fft = None
    if tracer_type == "abc":
        SN0 = SN1 = SN2 = SN3 = 0.0
    else:
        SN0 = N_gal_a if (correlation_mode == "auto" and [ell_1, ell_2, L] == [0, 0, 0]) else 0.0
        P_field = ca * np.conj(ca if correlation_mode == "auto" else cb) * boxsize.prod() ** 2
        if correlation_mode == "auto":
            P_field = P_field - A.compensate_bk_noise(w_circ, np.full(P_field.shape, SN0, dtype=np.complex128), sampler)
        if L != 0:
            P_field = P_field * A.get_Ylm(L, 0, True)(*kgrid)
        total_sum = A.radial_binning(P_field, k3_bins, k3_edge, knorm)
        SN1, SN2 = 0.0, 0.0
        if ell_1 == L and ell_2 == 0:
            if correlation_mode == "auto":
                SN1 = (total_sum / k_num_k3)[: len(k_center)] * (2 * L + 1)
                if ell_1 == 0:
                    SN2 = SN1.copy()
            elif tracer_type == "abb":
                SN1 = (total_sum / k_num_k3)[: len(k_center)] * (2 * L + 1)
        if dvm == "full":
            S1 = np.zeros((k_bins, k_bins)).astype("c16")
            S2 = np.zeros((k_bins, k_bins)).astype("c16")
            if np.ndim(SN1) == 0:
                S1[:] = SN1
            else:
                for i in range(k_bins):
                    S1[i, :] = SN1[i]
            if np.ndim(SN2) == 0:
                S2[:] = SN2
            else:
                for i in range(k_bins):
                    S2[:, i] = SN2[i]
            SN1, SN2 = S1, S2
        if sn_mode == "ana":
            SN3 = s3_ana(total_sum)
        elif sn_mode == "fft":
            SN3 = s3_fft()
        elif sn_mode == "both":
            SN3 = s3_ana(total_sum)
            SN3_fft = s3_fft()
        else:
            raise ValueError("shotnoise_mode must be either 'ana', 'fft', or 'both'.")

    # normalisation (``:1495-1548``)
    V = boxsize.prod()
    if correlation_mode == "auto":
        I_norm = (N_gal_a ** 3) / V ** 2
    elif tracer_type == "aab":
        I_norm = (N_gal_a ** 2 * N_gal_b) / V ** 2
    elif tracer_type == "abb":
        I_norm = (N_gal_a * N_gal_b ** 2) / V ** 2
    else:
        I_norm = (N_gal_a * N_gal_b * N_gal_c) / V ** 2
    total_sn = SN0 + SN1 + SN2 + SN3
    total_res = total_res * (N_ells * H_ells / I_norm)
    if dvm == "full":
        total_res *= V ** 2 / np.outer(k_num[:k_bins], k_num[:k_bins])
    else:
        total_res *= V ** 2 / (k_num[: len(total_res)]) ** 2
    total_res *= vpc
    total_sn = total_sn * (1 / I_norm)
    sn_terms = {"SN0": SN0 / I_norm, "SN1": SN1 / I_norm, "SN2": SN2 / I_norm, "SN3": SN3 / I_norm}
    if sn_mode == "both":
        sn_terms["SN3_FFT"] = SN3_fft / I_norm
    return {
        "B_sugi": total_res - total_sn,
        "SN_terms": sn_terms,
        "I_norm": I_norm,
        "Shot_noise": total_sn,
        "Bk_raw": total_res,
        "k_eff": k_eff,
        "nmodes": k_num,
    }
