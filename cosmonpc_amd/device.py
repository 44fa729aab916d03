"""Device-side handle: one cnpc context per (device, stream), torch tensors
for every buffer.  Thin typed wrappers over the C ABI (include/cnpc.h); every
call validates shapes and dtypes on the host before anything is launched.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, cnpc_poly

WINDOW = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}
PREC = {"fp64": 0, "fp32": 1}
REAL_DTYPE = {0: torch.float64, 1: torch.float32}
CPLX_DTYPE = {0: torch.complex128, 1: torch.complex64}
GEMM_MODE = {"bf16x3": 0, "fp16x3": 1, "gemm_fp64": 2, "fp8x6": 3}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class GpuContext:
    """Owns a ``cnpc_ctx`` bound to torch's current stream on ``device``."""

    def __init__(self, device=None):
        if not torch.cuda.is_available():
            raise _lib.CnpcError("no GPU visible to torch: the MI355X path has no CPU fallback")
        self.lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        torch.cuda.set_device(self.device)
        h = ctypes.c_void_p()
        check(self.lib.cnpc_ctx_create(ctypes.byref(h), self.device.index), "cnpc_ctx_create")
        self.h = h
        self.bind_stream()

    def bind_stream(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(self.lib.cnpc_ctx_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)), "cnpc_ctx_set_stream")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.cnpc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- buffers
    def mesh(self, n, prec, batch=None):
        shape = (n, n, n) if batch is None else (batch, n, n, n)
        return torch.empty(shape, dtype=REAL_DTYPE[prec], device=self.device)

    def spectrum(self, n, prec, batch=None):
        shape = (n, n, n // 2 + 1) if batch is None else (batch, n, n, n // 2 + 1)
        return torch.empty(shape, dtype=CPLX_DTYPE[prec], device=self.device)

    @staticmethod
    def _check_mesh(t, n, prec, batch=None):
        shape = (n, n, n) if batch is None else (batch, n, n, n)
        if t.dtype != REAL_DTYPE[prec] or tuple(t.shape) != shape or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected contiguous device mesh {shape} {REAL_DTYPE[prec]}, got {tuple(t.shape)} {t.dtype}")

    @staticmethod
    def _check_spec(t, n, prec, batch=None):
        shape = (n, n, n // 2 + 1) if batch is None else (batch, n, n, n // 2 + 1)
        if t.dtype != CPLX_DTYPE[prec] or tuple(t.shape) != shape or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected contiguous device spectrum {shape} {CPLX_DTYPE[prec]}, got {tuple(t.shape)} {t.dtype}")

    # ---------------------------------------------------------------- kernels
    def paint(self, pos, w, n, boxsize, sampler, interlaced, prec, mesh0, mesh1=None, boxcenter=None):
        if pos.dtype != torch.float64 or w.dtype != torch.float64:
            raise ValueError("positions and weights must be float64")
        if pos.dim() != 2 or pos.shape[1] != 3 or w.shape != (pos.shape[0],):
            raise ValueError("positions must be (Np, 3) and weights (Np,)")
        if not (pos.is_contiguous() and w.is_contiguous() and pos.is_cuda and w.is_cuda):
            raise ValueError("positions / weights must be contiguous device tensors")
        self._check_mesh(mesh0, n, prec)
        if interlaced:
            self._check_mesh(mesh1, n, prec)
        center = None
        if boxcenter is not None:
            center = (ctypes.c_double * 3)(*[float(c) for c in boxcenter])
        check(self.lib.cnpc_paint(self.h, _ptr(pos), _ptr(w), pos.shape[0], n, float(boxsize), WINDOW[sampler],
                                  int(bool(interlaced)), center, prec, _ptr(mesh0), _ptr(mesh1)), "cnpc_paint")

    def paint_batch(self, pos, w, offsets, n, boxsize, sampler, interlaced, prec, mesh0, mesh1=None, off_dev=None,
                    bounds=None):
        """Paint the catalogs [offsets[b], offsets[b+1]) of the concatenated
        device arrays into meshes mesh0[b] (and mesh1[b]); ``offsets`` is a
        host int sequence of nbatch + 1 entries.  For a call recorded in a
        graph and replayed with other catalogs: ``off_dev`` is a device int64
        tensor holding the offsets (rewritten before each replay), ``pos``
        and ``w`` are capacity-sized buffers, and ``bounds`` = (total, max)
        upper bounds of the particle count and the largest catalog."""
        off = np.asarray(offsets, dtype=np.int64)
        nb = len(off) - 1
        if nb < 1 or off[0] != 0 or np.any(np.diff(off) < 0):
            raise ValueError("offsets must start at 0 and increase")
        if bounds is not None:
            if off_dev is None or off_dev.dtype != torch.int64 or off_dev.numel() != nb + 1 or not off_dev.is_cuda:
                raise ValueError("bounds need the device offsets (int64, nbatch + 1 entries)")
            total, biggest = int(bounds[0]), int(bounds[1])
            if total > pos.shape[0] or total < off[-1] or biggest < int(np.diff(off).max()) or biggest > total:
                raise ValueError("bounds must cover the catalogs and fit the particle buffers")
        elif off[-1] != pos.shape[0]:
            raise ValueError("offsets must end at the particle count")
        if (pos.dtype != torch.float64 or w.dtype != torch.float64 or pos.dim() != 2 or pos.shape[1] != 3
                or w.shape != (pos.shape[0],) or not (pos.is_contiguous() and w.is_contiguous() and pos.is_cuda)):
            raise ValueError("positions must be a contiguous (Np, 3) and weights an (Np,) float64 device tensor")
        self._check_mesh(mesh0, n, prec, nb)
        if interlaced:
            self._check_mesh(mesh1, n, prec, nb)
        if bounds is None:
            off_d = torch.from_numpy(off).to(self.device)
            total, biggest = int(off[-1]), int(np.diff(off).max())
        else:
            off_d = off_dev
        check(self.lib.cnpc_paint_batch(self.h, _ptr(pos), _ptr(w), _ptr(off_d), nb, total, biggest, n,
                                        float(boxsize), WINDOW[sampler], int(bool(interlaced)), None, prec,
                                        _ptr(mesh0), _ptr(mesh1 if interlaced else None)), "cnpc_paint_batch")

    def r2c(self, mesh, out, n, prec, batch=None):
        self._check_mesh(mesh, n, prec, batch)
        self._check_spec(out, n, prec, batch)
        check(self.lib.cnpc_fft_r2c(self.h, n, batch or 1, prec, _ptr(mesh), _ptr(out)), "cnpc_fft_r2c")

    def c2r(self, spec, out, n, prec, batch=None):
        """Inverse FFT; ``spec`` is overwritten."""
        self._check_spec(spec, n, prec, batch)
        self._check_mesh(out, n, prec, batch)
        check(self.lib.cnpc_fft_c2r(self.h, n, batch or 1, prec, _ptr(spec), _ptr(out)), "cnpc_fft_c2r")

    def prepare(self, c1, c2, out, n, boxsize, prec, interlaced, comp, sampler, scale, sub0=0.0, zero_k0=False):
        self._check_spec(c1, n, prec)
        self._check_spec(out, n, prec)
        if interlaced:
            self._check_spec(c2, n, prec)
        check(self.lib.cnpc_prepare_spectrum(self.h, n, float(boxsize), prec, _ptr(c1), _ptr(c2 if interlaced else None),
                                             int(bool(interlaced)), int(comp), WINDOW[sampler], float(scale),
                                             float(sub0), int(bool(zero_k0)), _ptr(out)), "cnpc_prepare_spectrum")

    def bin(self, a, b, n, boxsize, prec, edges, polys=(), ab_scale=1.0, noise_window=0, noise_amp=0.0):
        """Returns host arrays (count, ksum, re[npoly, nb], im[npoly, nb])."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        nb = len(edges) - 1
        npoly = len(polys)
        if npoly:
            self._check_spec(a, n, prec)
            self._check_spec(b, n, prec)
        out = self.bin_dev(a, b, n, boxsize, prec, edges, polys, ab_scale, noise_window, noise_amp)
        return self.split_bins(out.cpu().numpy(), nb, npoly)

    def bin_dev(self, a, b, n, boxsize, prec, edges, polys=(), ab_scale=1.0, noise_window=0, noise_amp=0.0):
        """``bin`` without the read-back: the device FP64 vector [count, ksum,
        re (npoly x nb), im (npoly x nb)] (split it with ``split_bins``)."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        nb = len(edges) - 1
        npoly = len(polys)
        if npoly:
            self._check_spec(a, n, prec)
            self._check_spec(b, n, prec)
        out = torch.empty((2 + 2 * npoly) * nb, dtype=torch.float64, device=self.device)
        parr = (cnpc_poly * max(npoly, 1))(*polys) if npoly else None
        check(self.lib.cnpc_bin_spectrum(self.h, n, float(boxsize), prec, _ptr(a if npoly else None),
                                         _ptr(b if npoly else None), edges.ctypes.data_as(ctypes.c_void_p), nb, npoly,
                                         ctypes.cast(parr, ctypes.c_void_p) if parr is not None else None,
                                         float(ab_scale), int(noise_window), float(noise_amp), _ptr(out)),
              "cnpc_bin_spectrum")
        return out

    @staticmethod
    def split_bins(h, nb, npoly):
        """(count, ksum, re[npoly, nb], im[npoly, nb]) of a host copy of a
        ``bin_dev`` vector."""
        return h[:nb].copy(), h[nb:2 * nb].copy(), h[2 * nb:(2 + npoly) * nb].reshape(npoly, nb), \
            h[(2 + npoly) * nb:].reshape(npoly, nb)

    def cross_spectrum(self, a, b, out, n, prec, ab_scale, noise_window, noise_amp):
        for t in (a, b, out):
            self._check_spec(t, n, prec)
        check(self.lib.cnpc_cross_spectrum(self.h, n, prec, _ptr(a), _ptr(b), float(ab_scale), int(noise_window),
                                           float(noise_amp), _ptr(out)), "cnpc_cross_spectrum")

    def resample(self, spec, n, m, prec, nmax, out=None):
        """m-grid half spectrum holding the |n_d| <= nmax modes of ``spec``."""
        self._check_spec(spec, n, prec)
        if out is None:
            out = self.spectrum(m, prec)
        self._check_spec(out, m, prec)
        check(self.lib.cnpc_resample_spectrum(self.h, n, m, prec, _ptr(spec), int(nmax), _ptr(out)),
              "cnpc_resample_spectrum")
        return out

    def shell_synth(self, delta, n, boxsize, prec, edges, bin0, nshell, poly, work, out):
        """Real shells out[s] (nshell, n, n, n) by the pruned synthesis;
        ``work`` holds nshell half-spectrum planes stacks."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        if delta is not None:
            self._check_spec(delta, n, prec)
        self._check_spec(work[:nshell], n, prec, nshell)
        self._check_mesh(out, n, prec, nshell)
        check(self.lib.cnpc_shell_synth(self.h, n, float(boxsize), prec, _ptr(delta),
                                        edges.ctypes.data_as(ctypes.c_void_p), len(edges) - 1, int(bin0), int(nshell),
                                        ctypes.byref(poly), _ptr(work), _ptr(out)), "cnpc_shell_synth")

    @staticmethod
    def synth_radius(boxsize, kmax):
        """R = ceil(kmax / kf): every shell mode has |n_d| <= R."""
        return int(math.ceil(float(kmax) / (2.0 * math.pi / float(boxsize))))

    @staticmethod
    def half_layout(n, R, nslots):
        """Shape of the half-synthesised shells buffer for grid n and radius
        R, and the axis holding the slots.  Blocked
        [x][slot][kz/8][y/2][kz%8][y%2] for the n = 400 plans when 4 (R+1)
        <= n, else tile-major [x][y/2][slot][kz][y%2] (kz < R+1)."""
        Q = R + 1
        if n == 400 and 4 * Q <= n:
            return (n, nslots, (Q + 7) // 8, n // 2, 8, 2), 1
        return (n, n // 2, nslots, Q, 2), 2

    def half_shells(self, n, R, nslots):
        """Buffer of the half-synthesised shells (complex64, ``half_layout``)."""
        return torch.empty(self.half_layout(n, R, nslots)[0], dtype=torch.complex64, device=self.device)

    def _half_slots(self, H, n, R):
        """Slot count of a half-shells buffer; raises if H is not a
        contiguous complex64 device buffer of ``half_layout(n, R, .)``."""
        if H.dtype == torch.complex64 and H.is_contiguous() and H.is_cuda:
            shape, axis = self.half_layout(n, R, 1)
            nslots = H.shape[axis]
            if tuple(H.shape) == self.half_layout(n, R, nslots)[0]:
                return nslots
        raise ValueError(f"half shells must be a contiguous complex64 tensor of shape "
                         f"{self.half_layout(n, R, 'nslots')[0]}")

    def shell_synth_half(self, delta, n, boxsize, edges, bin0, nshell, poly, work, out, slot0, reuse_cube=False):
        """x and y passes of the FP32 shell synthesis of bins bin0 .. bin0 +
        nshell - 1 into slots slot0 .. of ``out`` (from ``half_shells``).
        ``reuse_cube``: same delta, poly and edges as the previous synthesis
        call (the mode cube is kept).  Returns the kz rows written per shell
        (numpy int32)."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        R = self.synth_radius(boxsize, edges[-1])
        if delta is not None:
            self._check_spec(delta, n, 1)
        self._check_spec(work[:nshell], n, 1, nshell)
        nslots = self._half_slots(out, n, R)
        rows = np.zeros(int(nshell), dtype=np.int32)
        check(self.lib.cnpc_shell_synth_half(self.h, n, float(boxsize), _ptr(delta),
                                             edges.ctypes.data_as(ctypes.c_void_p), len(edges) - 1, int(bin0),
                                             int(nshell), ctypes.byref(poly), _ptr(work), _ptr(out),
                                             nslots, int(slot0), rows.ctypes.data_as(ctypes.c_void_p),
                                             int(bool(reuse_cube))),
              "cnpc_shell_synth_half")
        return rows

    def zgemm_supported(self, n, na, nb, same):
        ok = ctypes.c_int(0)
        check(self.lib.cnpc_contract_zgemm_supported(self.h, n, int(na), int(nb), int(bool(same)), ctypes.byref(ok)),
              "cnpc_contract_zgemm_supported")
        return bool(ok.value)

    def contract_zgemm(self, n, boxsize, kmax, G, HA, na, kz_a, HB=None, nb=None, kz_b=None, nblk=0,
                       with_bb=False, noalloc=False, out=None):
        """C[a, b] = sum_x A_a G B_b over the shells in slots 0 .. na-1 of HA
        and 0 .. nb-1 of HB (None: B = A), z pass fused (device float64
        tensor).  kz_a / kz_b: kz rows per slot as ``shell_synth_half``
        returned them (the rows above are not written by the synthesis).
        ``with_bb`` (B != A): returns (C, C_bb) with C_bb[b, c] = sum_x B_b G
        B_c from the same z transforms.
        ``noalloc``: through cnpc_contract_zgemm_noalloc (partials in the
        library's static scratch; nothing allocated by the call).  ``out``:
        caller-allocated C (and C_bb) tensors to write into."""
        if kz_a is None or (HB is not None and kz_b is None):
            raise ValueError("contract_zgemm needs the kz rows of every slot")
        if with_bb and (HB is None or HB is HA):
            raise ValueError("contract_zgemm: with_bb needs a B different from A")
        if G.dtype != torch.float32 or G.numel() != n ** 3 or not G.is_contiguous():
            raise ValueError("G must be a contiguous FP32 n^3 mesh")
        R = self.synth_radius(boxsize, kmax)
        sa = self._half_slots(HA, n, R)
        sb = sa if HB is None else self._half_slots(HB, n, R)
        nb_ = na if HB is None else nb

        def rows(kz, count):
            if kz is None:
                return None
            kz = np.ascontiguousarray(kz, dtype=np.int32)
            if len(kz) < count:
                raise ValueError("kz rows must cover every slot used")
            return kz

        ka, kb = rows(kz_a, na), rows(kz_b, nb_) if HB is not None else None
        if out is None:
            C = torch.empty((na, nb_), dtype=torch.float64, device=self.device)
            Cbb = torch.empty((nb_, nb_), dtype=torch.float64, device=self.device) if with_bb else None
        else:
            C, Cbb = (out if with_bb else (out, None))
            if (C.dtype != torch.float64 or tuple(C.shape) != (na, nb_) or not C.is_contiguous()
                    or (with_bb and (Cbb.dtype != torch.float64 or tuple(Cbb.shape) != (nb_, nb_)
                                     or not Cbb.is_contiguous()))):
                raise ValueError("out must be contiguous float64 (na, nb) [and (nb, nb)] tensors")
        fn = self.lib.cnpc_contract_zgemm_noalloc if noalloc else self.lib.cnpc_contract_zgemm
        check(fn(self.h, n, float(boxsize), float(kmax), _ptr(G), _ptr(HA), int(na), sa,
                 None if ka is None else ka.ctypes.data_as(ctypes.c_void_p), _ptr(HB), int(nb_), sb,
                 None if kb is None else kb.ctypes.data_as(ctypes.c_void_p), int(nblk), _ptr(C), _ptr(Cbb)),
              "cnpc_contract_zgemm_noalloc" if noalloc else "cnpc_contract_zgemm")
        return (C, Cbb) if with_bb else C

    def shell_filter(self, delta, n, boxsize, prec, edges, bin0, nshell, poly, out):
        """``delta`` None filters the weight alone."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        if delta is not None:
            self._check_spec(delta, n, prec)
        if out.dim() != 4 or out.shape[0] < nshell:
            raise ValueError("shell output must hold nshell spectra")
        self._check_spec(out[:nshell], n, prec, nshell)
        check(self.lib.cnpc_shell_filter(self.h, n, float(boxsize), prec, _ptr(delta),
                                         edges.ctypes.data_as(ctypes.c_void_p), len(edges) - 1, int(bin0), int(nshell),
                                         ctypes.byref(poly), _ptr(out)), "cnpc_shell_filter")

    def contract_pairs(self, prec, G, A, B, ia, ib):
        """sum_x G A[ia] B[ib] per pair, host float64 array."""
        ncell = G.numel()
        if A.dim() != 2 or B.dim() != 2 or A.shape[1] != ncell or B.shape[1] != ncell:
            raise ValueError("A / B must be (nshell, ncell) stacks matching G")
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ib = np.ascontiguousarray(ib, dtype=np.int32)
        if len(ia) != len(ib) or (len(ia) and (ia.max() >= A.shape[0] or ib.max() >= B.shape[0])):
            raise ValueError("pair index out of range")
        out = np.zeros(len(ia), dtype=np.float64)
        check(self.lib.cnpc_contract_pairs(self.h, prec, ncell, _ptr(G), _ptr(A), _ptr(B), len(ia),
                                           ia.ctypes.data_as(ctypes.c_void_p), ib.ctypes.data_as(ctypes.c_void_p),
                                           out.ctypes.data_as(ctypes.c_void_p)), "cnpc_contract_pairs")
        return out

    def contract_gemm(self, mode, G, A, B=None, sa=1.0, sb=1.0, nblk=0):
        """C[a, b] = sum_x A[a] G B[b] (device float64 tensor)."""
        same = B is None
        if same:
            B = A
        ncell = G.numel()
        if A.dim() != 2 or B.dim() != 2 or A.shape[1] != ncell or B.shape[1] != ncell:
            raise ValueError("A / B must be (nshell, ncell) stacks matching G")
        want = torch.float64 if GEMM_MODE[mode] == 2 else torch.float32
        for t in (G, A, B):
            if t.dtype != want or not t.is_contiguous():
                raise ValueError(f"{mode} needs contiguous {want} operands")
        C = torch.empty((A.shape[0], B.shape[0]), dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_contract_gemm(self.h, GEMM_MODE[mode], ncell, _ptr(G), _ptr(A), A.shape[0], _ptr(B),
                                          B.shape[0], int(same), float(sa), float(sb), int(nblk), _ptr(C)),
              "cnpc_contract_gemm")
        return C

    def absmax(self, x):
        out = ctypes.c_double()
        check(self.lib.cnpc_absmax_f32(self.h, _ptr(x), x.numel(), ctypes.byref(out)), "cnpc_absmax_f32")
        return out.value

    def sync(self):
        check(self.lib.cnpc_sync(self.h), "cnpc_sync")

    def reserve(self, nbytes):
        """Pre-size the context's reduction scratch (``cnpc_ctx_reserve``)."""
        check(self.lib.cnpc_ctx_reserve(self.h, int(nbytes)), "cnpc_ctx_reserve")

    # ---------------------------------------------------------------- survey (full grids)
    def full(self, n, prec, batch=None):
        shape = (n, n, n) if batch is None else (batch, n, n, n)
        return torch.empty(shape, dtype=CPLX_DTYPE[prec], device=self.device)

    @staticmethod
    def _check_full(t, n, prec):
        if t.dtype != CPLX_DTYPE[prec] or tuple(t.shape) != (n, n, n) or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected contiguous device full mesh {(n, n, n)} {CPLX_DTYPE[prec]}, "
                             f"got {tuple(t.shape)} {t.dtype}")

    def fft_c2c(self, src, out, n, prec, direction):
        """Unnormalised c2c transform of a full mesh; direction -1 forward, +1 inverse."""
        self._check_full(src, n, prec)
        self._check_full(out, n, prec)
        check(self.lib.cnpc_fft_c2c(self.h, n, 1, prec, int(direction), _ptr(src), _ptr(out)), "cnpc_fft_c2c")

    def full_spectrum(self, c1, c2, out, n, boxsize, prec, interlaced, comp, sampler, scale):
        self._check_spec(c1, n, prec)
        if interlaced:
            self._check_spec(c2, n, prec)
        self._check_full(out, n, prec)
        check(self.lib.cnpc_full_spectrum(self.h, n, float(boxsize), prec, _ptr(c1), _ptr(c2 if interlaced else None),
                                          int(bool(interlaced)), int(comp), WINDOW[sampler], float(scale), _ptr(out)),
              "cnpc_full_spectrum")

    def xweight(self, F, out, n, boxsize, prec, boxcenter, poly):
        self._check_full(F, n, prec)
        self._check_full(out, n, prec)
        c = (ctypes.c_double * 3)(*[float(v) for v in boxcenter])
        check(self.lib.cnpc_xweight(self.h, n, float(boxsize), prec, c, ctypes.byref(poly), _ptr(F), _ptr(out)),
              "cnpc_xweight")

    def kcombine(self, A, out, n, boxsize, prec, B=None, opb=0, comp=0, sampler="cic", shot=None, poly=None,
                 coef=1.0, accumulate=False):
        """out (+)= coef * w(khat) * [comp(k) (A op B) - shot * C_bk(k)]."""
        self._check_full(A, n, prec)
        self._check_full(out, n, prec)
        if opb:
            self._check_full(B, n, prec)
        s = complex(shot) if shot is not None else 0j
        c = complex(coef)
        check(self.lib.cnpc_kcombine(self.h, n, float(boxsize), prec, _ptr(A), _ptr(B if opb else None), int(opb),
                                     int(comp), WINDOW[sampler], int(shot is not None), s.real, s.imag,
                                     ctypes.byref(poly) if poly is not None else None, c.real, c.imag,
                                     int(bool(accumulate)), _ptr(out)), "cnpc_kcombine")

    def full_bin(self, A, n, boxsize, prec, edges, B=None, opb=0):
        """Complex per-bin sums of A op B over the full grid (host complex128)."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        nb = len(edges) - 1
        self._check_full(A, n, prec)
        if opb:
            self._check_full(B, n, prec)
        out = torch.empty(2 * nb, dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_full_bin(self.h, n, float(boxsize), prec, _ptr(A), _ptr(B if opb else None), int(opb),
                                     edges.ctypes.data_as(ctypes.c_void_p), nb, _ptr(out)), "cnpc_full_bin")
        o = out.cpu().numpy()
        return o[:nb] + 1j * o[nb:]

    def triple_sum(self, prec, A, B=None, C=None):
        """sum_x A*B*C over complex meshes of equal size (host complex)."""
        want = CPLX_DTYPE[prec]
        for t in (A, B, C):
            if t is not None and (t.dtype != want or not t.is_contiguous() or not t.is_cuda or t.numel() != A.numel()):
                raise ValueError("triple_sum operands must be contiguous device complex meshes of one size")
        out = torch.empty(2, dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_triple_sum(self.h, prec, A.numel(), _ptr(A), _ptr(B), _ptr(C), _ptr(out)),
              "cnpc_triple_sum")
        o = out.cpu().numpy()
        return complex(o[0], o[1])

    def split_complex(self, src, prec):
        if src.dtype != CPLX_DTYPE[prec] or not src.is_contiguous() or not src.is_cuda:
            raise ValueError("split_complex needs a contiguous device complex tensor")
        re = torch.empty(src.shape, dtype=REAL_DTYPE[prec], device=self.device)
        im = torch.empty(src.shape, dtype=REAL_DTYPE[prec], device=self.device)
        check(self.lib.cnpc_split_complex(self.h, prec, src.numel(), _ptr(src), _ptr(re), _ptr(im)),
              "cnpc_split_complex")
        return re, im

    def resample_full(self, src, n, m, prec, nmax, out=None):
        self._check_full(src, n, prec)
        if out is None:
            out = self.full(m, prec)
        self._check_full(out, m, prec)
        check(self.lib.cnpc_resample_full(self.h, n, m, prec, _ptr(src), int(nmax), _ptr(out)), "cnpc_resample_full")
        return out

    def reflect_conj(self, A, n, prec, c1=0.5, c2=0.5):
        """c1 A(k) + c2 conj(A(-k)) as a new full spectrum."""
        self._check_full(A, n, prec)
        out = self.full(n, prec)
        check(self.lib.cnpc_reflect_conj(self.h, n, prec, _ptr(A), float(c1), float(c2), _ptr(out)),
              "cnpc_reflect_conj")
        return out

    @staticmethod
    def _check_vec(t, np_, what):
        if t is None:
            return
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda or t.shape != (np_,):
            raise ValueError(f"{what} must be a contiguous device float64 vector of length {np_}")

    def particle_sums(self, pos, W, Wfkp, NZ=None, polys=()):
        """FP64 catalog sums (host array): [sum W Wfkp, sum W, sum W NZ Wfkp^2,
        sum (W Wfkp)^2, sum NZ^2 W^2 Wfkp^3, then (re, im) per poly]."""
        np_ = W.shape[0]
        self._check_vec(W, np_, "W")
        self._check_vec(Wfkp, np_, "WEIGHT_FKP")
        self._check_vec(NZ, np_, "NZ")
        if polys and (pos.dtype != torch.float64 or tuple(pos.shape) != (np_, 3) or not pos.is_contiguous()):
            raise ValueError("positions must be a contiguous (Np, 3) float64 device tensor")
        npoly = len(polys)
        parr = (cnpc_poly * max(npoly, 1))(*polys) if npoly else None
        out = torch.empty(5 + 2 * npoly, dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_particle_sums(self.h, np_, _ptr(pos if npoly else None), _ptr(W), _ptr(Wfkp), _ptr(NZ),
                                          npoly, ctypes.cast(parr, ctypes.c_void_p) if parr is not None else None,
                                          _ptr(out)), "cnpc_particle_sums")
        return out.cpu().numpy()

    def pair_weights(self, W, Wfkp, power=1, scale=1.0):
        np_ = W.shape[0]
        self._check_vec(W, np_, "W")
        self._check_vec(Wfkp, np_, "WEIGHT_FKP")
        out = torch.empty(np_, dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_pair_weights(self.h, np_, _ptr(W), _ptr(Wfkp), int(power), float(scale), _ptr(out)),
              "cnpc_pair_weights")
        return out

    def radecz_to_xyz_flrw(self, ra, dec, z, cosmo, dh):
        """RA/DEC/Z -> positions in a flat FLRW cosmology with radiation and
        neutrinos (``cosmo``: the 9 numbers of cnpc_radecz_to_xyz_flrw,
        ``dh`` = c/H0 in the output unit)."""
        np_ = ra.shape[0]
        for t, nm in ((ra, "RA"), (dec, "DEC"), (z, "Z")):
            self._check_vec(t, np_, nm)
        c = np.ascontiguousarray(cosmo, dtype=np.float64)
        if c.shape != (9,):
            raise ValueError("cosmo must hold 9 numbers")
        out = torch.empty((np_, 3), dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_radecz_to_xyz_flrw(self.h, np_, _ptr(ra), _ptr(dec), _ptr(z),
                                               c.ctypes.data_as(ctypes.c_void_p), float(dh), _ptr(out)),
              "cnpc_radecz_to_xyz_flrw")
        return out

    def radecz_to_xyz(self, ra, dec, z, omega_m):
        np_ = ra.shape[0]
        for t, nm in ((ra, "RA"), (dec, "DEC"), (z, "Z")):
            self._check_vec(t, np_, nm)
        out = torch.empty((np_, 3), dtype=torch.float64, device=self.device)
        check(self.lib.cnpc_radecz_to_xyz(self.h, np_, _ptr(ra), _ptr(dec), _ptr(z), float(omega_m), _ptr(out)),
              "cnpc_radecz_to_xyz")
        return out


_CTX = {}


def get_context(device=None):
    """Process-wide context per device (created lazily)."""
    idx = torch.cuda.current_device() if device is None else device
    if idx not in _CTX:
        _CTX[idx] = GpuContext(idx)
    ctx = _CTX[idx]
    ctx.bind_stream()
    return ctx
