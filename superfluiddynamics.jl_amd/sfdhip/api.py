"""Host-side mirror of the reference's Grid / Field / NumModel API for the
split-step path, backed by libsfdhip.so.

Names, argument meaning and return values follow the reference:

=====================================  =========================================
reference (file:line)                  here
=====================================  =========================================
``Grid(size, bounds; FT, array_type)``  ``Grid(size, bounds, FT=..., array_type=HipArray)``
  (Grids.jl:64-129)
``Field(grid, ComplexField())``         ``Field(grid, ComplexField())``  (Fields.jl:94-157)
``norm(f)``, ``normalize!(f)``          ``norm(f)``, ``normalize_(f)``  (Fields.jl:172-200)
``GrossPitaevskiiParameters(...)``      same keywords (Parameters.jl:49-55)
``PotentialZero``, ``PotentialQuadratic``  same (Potentials.jl:17-25, quadratic.jl:48-67)
``NumModelSplit1/2(f, p, Δt, n, fb)``  same positional arguments (splitting.jl:58-69,129-140)
``timeStep!(n)`` -> 1 / 2               ``timeStep(n)`` -> 1 / 2  (splitting.jl:78-82,149-154)
``energy(n, show)``                     ``energy(n, show)`` -> (EΩ, EΔ, Eβ, E) (GrossPitaevskii.jl:61-136)
``solve!(n; istart)``                   ``solve(n, istart=1)`` (NumModels.jl:23-46, no writers)
``mul_x!/ldiv_x!`` ... ``_z!``          ``mul_x/ldiv_x`` ... (Plans.jl:1132-1250)
=====================================  =========================================

Storage: a field is a torch tensor on the current HIP device, complex64
(FT = float32) or complex128 (FT = float64), shape ``(nz, ny, nx)`` /
``(ny, nx)`` in C order -- the reference's column-major ``(nx, ny[, nz])``
pen_x layout, x fastest.  Geometry stays in FP64 whatever FT is.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


class HipArray:
    """The ``array_type`` plug of Grid: fields live in HIP device memory."""


class ComplexField:
    pass


class RealField:
    pass


def _require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise L.SfdError(L.SFD_ERR_HIP, "no HIP device is available; the split-step path runs on the GPU only")


def linrange_coords(lo: float, hi: float, n: int) -> np.ndarray:
    """``LinRange(lo, hi, n+1)[1:n]`` with Julia's ``(1-t) lo + t hi``, t = i/n
    (Grids.jl:78-82)."""
    t = np.arange(n, dtype=np.float64) / float(n)
    return (1.0 - t) * float(lo) + t * float(hi)


def fftfreq(n: int, d: float) -> np.ndarray:
    """``fftfreq(n, 2π/Δ)`` (Plans.jl:829-830): k * ((2π/Δ)/n)."""
    mult = (6.283185307179586 / d) / n
    k = np.arange(n)
    k = np.where(k < (n + 1) // 2, k, k - n)
    return k.astype(np.float64) * mult


class Grid:
    """Grid{N,FT,A} (Grids.jl:24-129).  ``size`` and ``bounds`` are x-first."""

    def __init__(self, size, bounds, FT=np.float64, array_type=HipArray, compute=None):
        size = tuple(int(s) for s in size)
        if len(size) not in (2, 3) or len(bounds) != len(size):
            raise L.ArgError(L.SFD_ERR_ARG, "Grid: 2D or 3D size with one (min, max) per axis")
        for s in size:
            if not s > 0:
                raise L.ArgError(L.SFD_ERR_ARG, "Grid: sizes must be positive (Grids.jl:69)")
        self.n = size
        self.min = tuple(float(b[0]) for b in bounds)
        self.max = tuple(float(b[1]) for b in bounds)
        self.L = tuple(hi - lo for lo, hi in zip(self.min, self.max))
        for l in self.L:
            if not l > 0:
                raise L.ArgError(L.SFD_ERR_ARG, "Grid: lengths must be positive (Grids.jl:75)")
        self.Δ = tuple(l / s for l, s in zip(self.L, size))
        self.FT = np.dtype(FT).type
        if self.FT not in (np.float32, np.float64):
            raise L.ArgError(L.SFD_ERR_ARG, "Grid: FT must be Float32 or Float64")
        self.compute = np.dtype(compute).type if compute is not None else self.FT
        if self.compute not in (np.float32, np.float64) or (self.FT is np.float64 and self.compute is np.float32):
            raise L.ArgError(L.SFD_ERR_ARG, "Grid: compute must be FT or Float64 (FP32 fields with FP64 arithmetic)")
        if array_type is not HipArray:
            raise L.ArgError(L.SFD_ERR_ARG, "Grid: this backend provides array_type=HipArray only")
        self.array_type = array_type
        self.data = tuple(linrange_coords(lo, hi, s) for lo, hi, s in zip(self.min, self.max, size))

    ndims = property(lambda self: len(self.n))
    nx = property(lambda self: self.n[0])
    ny = property(lambda self: self.n[1])
    nz = property(lambda self: self.n[2])
    Lx = property(lambda self: self.L[0])
    Ly = property(lambda self: self.L[1])
    Lz = property(lambda self: self.L[2])
    Δx = property(lambda self: self.Δ[0])
    Δy = property(lambda self: self.Δ[1])
    Δz = property(lambda self: self.Δ[2])
    x = property(lambda self: self.data[0])
    y = property(lambda self: self.data[1])
    z = property(lambda self: self.data[2])

    @property
    def shape(self):
        """numpy/torch shape of a field: reversed extents (x fastest)."""
        return tuple(reversed(self.n))

    def __repr__(self):
        return (f"Grid{self.ndims}D\n  ├──────  resolution: {'×'.join(map(str, self.n))}\n"
                f"  └──────────  domain: " + "×".join(f"[{a},{b}]" for a, b in zip(self.min, self.max)))


class _Ctx:
    """Owns one sfd_ctx (the FFT plan state for a grid and precision)."""

    def __init__(self, grid: Grid):
        _require_gpu()
        lib = L.load()
        self.lib = lib
        nd = grid.ndims
        n = (ctypes.c_int64 * 3)(*(list(grid.n) + [1] * (3 - nd)))
        lo = (ctypes.c_double * 3)(*(list(grid.min) + [0.0] * (3 - nd)))
        hi = (ctypes.c_double * 3)(*(list(grid.max) + [1.0] * (3 - nd)))
        if grid.FT is np.float64:
            prec = L.SFD_F64
        else:
            prec = L.SFD_F32X if grid.compute is np.float64 else L.SFD_F32
        h = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        L.check(lib.sfd_ctx_create(ctypes.byref(h), nd, n, lo, hi, prec, ctypes.c_void_p(stream)))
        self.h = h

    def autotune(self, scratch=None):
        """Time every compiled sweep variant per axis and keep the fastest
        (sfd_autotune).  Returns {axis: [ms per launch of each variant]}."""
        times = (ctypes.c_float * 24)()
        ptr = ctypes.c_void_p(scratch.data_ptr()) if scratch is not None else None
        L.check(self.lib.sfd_autotune(self.h, ptr, times))
        return {a: [t for t in times[a * 8:(a + 1) * 8] if t >= 0] for a in range(3) if times[a * 8] >= 0}

    def variant(self, axis: int):
        """(index, count, (n1, n2, ts, g, pipelined, min waves/SIMD, halves)) of the
        sweep variant on axis."""
        idx, cnt = ctypes.c_int(), ctypes.c_int()
        shp = (ctypes.c_int * 7)()
        L.check(self.lib.sfd_ctx_get_variant(self.h, axis, ctypes.byref(idx), ctypes.byref(cnt), shp))
        return idx.value, cnt.value, tuple(shp)

    def set_variant(self, axis: int, index: int):
        L.check(self.lib.sfd_ctx_set_variant(self.h, axis, index))

    def sync_stream(self):
        L.check(self.lib.sfd_ctx_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.sfd_ctx_destroy(h)
            except Exception:
                pass
            self.h = None


class Field:
    """Field{N,1,FT,Complex{FT}} on HIP memory (Fields.jl:94-157)."""

    def __init__(self, g: Grid, t=None, ndims: int = 1):
        _require_gpu()
        if t is None:
            t = ComplexField()
        if ndims != 1:
            raise L.ArgError(L.SFD_ERR_ARG, "the split-step path uses scalar fields (ndims=1)")
        self.g = g
        complex_ = isinstance(t, ComplexField)
        if g.FT is np.float32:
            dt = torch.complex64 if complex_ else torch.float32
        else:
            dt = torch.complex128 if complex_ else torch.float64
        self.data = [torch.zeros(g.shape, dtype=dt, device="cuda")]
        self._ctx = None

    @property
    def ϕ(self):
        return self.data[0]

    @ϕ.setter
    def ϕ(self, value):
        self.data[0].copy_(torch.as_tensor(value).reshape(self.g.shape))

    phi = ϕ

    def ctx(self) -> _Ctx:
        if self._ctx is None:
            self._ctx = _Ctx(self.g)
        self._ctx.sync_stream()
        return self._ctx

    def autotune(self):
        """Pick the fastest sweep variant per axis for this field's grid,
        timed on a scratch field (ϕ is not touched)."""
        return self.ctx().autotune()

    def coords(self):
        """Broadcast-shaped FP64 numpy coordinates (x, y[, z]) like ``field.x``."""
        nd = self.g.ndims
        out = []
        for a in range(nd):
            shp = [1] * nd
            shp[nd - 1 - a] = self.g.n[a]
            out.append(self.g.data[a].reshape(shp))
        return tuple(out)

    x = property(lambda self: self.coords()[0])
    y = property(lambda self: self.coords()[1])
    z = property(lambda self: self.coords()[2])

    def set(self, values):
        """Copy a host array (numpy, x-fastest shape) into ϕ."""
        arr = np.ascontiguousarray(values)
        if not arr.flags.writeable:  # torch wants writable buffers (broadcast views are not)
            arr = arr.copy()
        self.data[0].copy_(torch.as_tensor(arr).to(self.data[0].dtype))

    def to_numpy(self):
        return self.data[0].detach().cpu().numpy()


def norm(f: Field) -> float:
    """norm(f) = sqrt(Σ|ϕ|²) sqrt(ΠΔ) with an FP64 sum (Fields.jl:172-180)."""
    c = f.ctx()
    s = ctypes.c_double()
    L.check(c.lib.sfd_norm2(c.h, ctypes.c_void_p(f.ϕ.data_ptr()), ctypes.byref(s)))
    return math.sqrt(s.value) * math.sqrt(math.prod(f.g.Δ))


def normalize_(f: Field) -> None:
    """normalize!(f) (Fields.jl:196-200)."""
    c = f.ctx()
    L.check(c.lib.sfd_normalize(c.h, ctypes.c_void_p(f.ϕ.data_ptr()), None))


normalize = normalize_


# ------------------------------------------------------------- parameters
class PotentialZero:
    """V = 0 (Potentials.jl:17-25); passed to the kernels as a NULL V."""

    def __init__(self, f: Field):
        self.f = f
        self._V = None

    @property
    def V(self):
        if self._V is None:
            self._V = torch.zeros(self.f.g.shape, dtype=_real_dtype(self.f), device="cuda")
        return self._V

    def device_ptr(self):
        return None


def _real_dtype(f: Field):
    return torch.float32 if f.g.FT is np.float32 else torch.float64


def _check_device_tensor(t, name: str, dtype, shape):
    """Raise unless ``t`` is a contiguous HIP-device tensor of ``dtype`` and
    ``shape``: its raw pointer goes straight to the kernels, which would
    read a host pointer, or past the end of a smaller or narrower buffer."""
    if not torch.is_tensor(t):
        raise L.ArgError(L.SFD_ERR_ARG, f"{name} must be a torch tensor on the HIP device")
    if t.device.type != "cuda":
        raise L.ArgError(L.SFD_ERR_ARG, f"{name} is on {t.device}; it must be on the HIP device")
    if t.dtype != dtype:
        raise L.ArgError(L.SFD_ERR_ARG, f"{name} has dtype {t.dtype}; {dtype} is required")
    if tuple(t.shape) != tuple(shape):
        raise L.ShapeError(L.SFD_ERR_SHAPE, f"{name} has shape {tuple(t.shape)}; {tuple(shape)} is required")
    if not t.is_contiguous():
        raise L.ArgError(L.SFD_ERR_ARG, f"{name} must be contiguous")


class PotentialQuadratic:
    """V = ½(1-α)(γx x² + γy y² [+ γz z²]) (quadratic.jl:61-67), evaluated in
    FP64 with the reference's operation order and rounded once to FT."""

    def __init__(self, f: Field, α=0.0, γx=1.0, γy=1.0, γz=1.0, **kw):
        α = kw.get("alpha", α)
        γx, γy, γz = kw.get("gx", γx), kw.get("gy", γy), kw.get("gz", γz)
        self.f, self.α, self.γx, self.γy, self.γz = f, float(α), float(γx), float(γy), float(γz)
        co = f.coords()
        if f.g.ndims == 2:
            x, y = co
            Vh = 0.5 * (1 - self.α) * (self.γx * x ** 2 + self.γy * y ** 2)
        else:
            x, y, z = co
            Vh = 0.5 * (1 - self.α) * (self.γx * x ** 2 + self.γy * y ** 2 + self.γz * z ** 2)
        Vh = np.broadcast_to(Vh, f.g.shape).astype(f.g.FT)
        self.V = torch.as_tensor(np.ascontiguousarray(Vh)).to("cuda")

    def device_ptr(self):
        return self.V.data_ptr()


class PotentialArray:
    """Any precomputed real V (same shape and FT as the field)."""

    def __init__(self, f: Field, V):
        V = torch.as_tensor(np.ascontiguousarray(V) if isinstance(V, np.ndarray) else V)
        self.f = f
        self.V = V.to(device="cuda", dtype=_real_dtype(f)).reshape(f.g.shape).contiguous()

    def device_ptr(self):
        return self.V.data_ptr()


class GrossPitaevskiiParameters:
    """GrossPitaevskiiParameters(; coeffΔ=-0.5, β=1, pot, Ω=0) (Parameters.jl:49-55)."""

    def __init__(self, *, pot, coeffΔ=-0.5, β=1.0, Ω=0.0, **kw):
        self.pot = pot
        self.coeffΔ = float(kw.get("coeffD", coeffΔ))
        self.β = float(kw.get("beta", β))
        self.Ω = float(kw.get("Omega", Ω))


def _vptr(param):
    p = param.pot.device_ptr() if param.pot is not None else None
    return ctypes.c_void_p(p) if p else None


# ----------------------------------------------------------------- models
class _SplitBase:
    order = 2
    mode = L.SFD_REAL_TIME

    def __init__(self, f: Field, param: GrossPitaevskiiParameters, Δt, niter, freqbckp):
        if param.pot is not None and getattr(param.pot, "f", f).g is not f.g:
            raise L.ArgError(L.SFD_ERR_ARG, "the potential belongs to another grid")
        self.f, self.param = f, param
        self.Δt, self.niter, self.freqbckp = float(Δt), int(niter), int(freqbckp)

    def steps(self, k: int) -> int:
        """k consecutive timeStep! calls in one enqueue; returns Σ of their returns."""
        c = self.f.ctx()
        p = self.param
        L.check(c.lib.sfd_split_step(c.h, ctypes.c_void_p(self.f.ϕ.data_ptr()), _vptr(p), p.coeffΔ, p.β, p.Ω,
                                     self.Δt, int(k), self.order, self.mode))
        return self.order * int(k)


class NumModelSplit1(_SplitBase):
    """Lie splitting K(Δt) N(Δt) (splitting.jl:58-82)."""
    order = 1

    def __repr__(self):
        return f"Splitting Order 1\n  ├───────  time step: {self.Δt}"


class NumModelSplit2(_SplitBase):
    """Strang splitting K(Δt/2) N(Δt) K(Δt/2) (splitting.jl:129-154)."""
    order = 2

    def __repr__(self):
        return f"Splitting Order 2\n  ├───────  time step: {self.Δt}"


class NumModelSplitImag(_SplitBase):
    """Imaginary-time split step (normalised gradient flow): the Split1/Split2
    sweeps with every exponent real, then normalize! after each step.  Its fixed
    point is the ground state that NumModelBackwardEuler reaches
    (backward-euler.jl:72-85) up to a splitting error: with Strang, O(Δt²)
    in the energy and O(Δt) in the state and μ (the nonlinear factor uses
    |ψ|² after the first half kinetic flow)."""
    mode = L.SFD_IMAG_TIME

    def __init__(self, f, param, Δt, niter, freqbckp, order: int = 2):
        super().__init__(f, param, Δt, niter, freqbckp)
        if order not in (1, 2):
            raise L.ArgError(L.SFD_ERR_ARG, "order must be 1 or 2")
        self.order = order


def timeStep(n: _SplitBase) -> int:
    """timeStep!(n): one split step in place; returns 1 (Split1) or 2 (Split2)."""
    return n.steps(1)


def solveLapRot(n: _SplitBase, Δtl: float) -> None:
    """solveLapRot!(n, Δtl) (splitting.jl:156-203): the linear flow over Δtl,
    x, y[, z] round trips (imaginary-time models use real exponents)."""
    c = n.f.ctx()
    p = n.param
    L.check(c.lib.sfd_lap_rot(c.h, ctypes.c_void_p(n.f.ϕ.data_ptr()), p.coeffΔ, p.Ω, float(Δtl), n.mode))


def solveNL(n: _SplitBase, Δtl: float) -> None:
    """solveNL!(n, Δtl) (splitting.jl:205-209): ϕ ← exp(-i(V + β|ϕ|²)Δtl) ϕ."""
    c = n.f.ctx()
    p = n.param
    L.check(c.lib.sfd_solve_nl(c.h, ctypes.c_void_p(n.f.ϕ.data_ptr()), _vptr(p), p.β, float(Δtl), n.mode))


def energy(n: _SplitBase, showEnergy: bool = False):
    """energy(n, showEnergy) -> (EΩ, EΔ, Eβ, E) (GrossPitaevskii.jl:61-136).
    The chemical potential is available as ``energy_all(n)[4]``."""
    EΩ, EΔ, Eβ, E, μ = energy_all(n)
    if showEnergy:
        if n.f.g.ndims == 2:
            print(f"Chemical potential : {μ}")
        print(f"Angular Momentum Energy : {EΩ}")
        print(f"Kinetic + Potential Energy : {EΔ}")
        print(f"Interaction Energy : {Eβ}")
        print(f"Total Energy: {E}")
    return EΩ, EΔ, Eβ, E


def energy_all(n: _SplitBase):
    """(EΩ, EΔ, Eβ, E, μ), μ = -EΩ + EΔ + 2Eβ (GrossPitaevskii.jl:79)."""
    c = n.f.ctx()
    p = n.param
    out = (ctypes.c_double * 5)()
    L.check(c.lib.sfd_energy(c.h, ctypes.c_void_p(n.f.ϕ.data_ptr()), _vptr(p), p.coeffΔ, p.β, p.Ω, out))
    return tuple(out)


def steps_with_energy(n: _SplitBase, k: int):
    """k iterations of (energy_all(n), timeStep(n)) in one call
    (sfd_split_step_energy): a (k, 5) array of (EΩ, EΔ, Eβ, E, μ) of the
    state before each step.  In real time with Ω = 0 the energies come out of
    the steps' own sweeps."""
    c = n.f.ctx()
    p = n.param
    out = np.zeros((int(k), 5))
    L.check(c.lib.sfd_split_step_energy(c.h, ctypes.c_void_p(n.f.ϕ.data_ptr()), _vptr(p), p.coeffΔ, p.β, p.Ω,
                                        n.Δt, int(k), n.order, n.mode,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def solve(n: _SplitBase, istart: int = 1, verbose: bool = False, writers=()):
    """solve!(n; istart) (NumModels.jl:23-46): niter+1 iterations of
    (energy, timeStep!); returns the history of (nstep, total, EΩ, EΔ, Eβ, E).
    ``writers`` (``sfdhip.io.WriterSave`` objects; the reference's VTK writer
    is not built) save "res-{istep}.h5" at istart and after every iteration
    it with it % freqbckp == 0.  The iterations between two backup points run
    in one sfd_split_step_energy call, with or without writers, so the
    history does not depend on the writers, and a run resumed from a backup
    with the same freqbckp repeats the original bit for bit (chained steps
    within a call order the commuting flows by the call's length); the
    printed lines follow each call."""
    from . import io as _io
    writers = list(writers)

    def save(istep):
        for w in writers:
            _io.write(w, prefix="res", istep=istep, Δt=n.Δt)

    save(istart)
    res = []
    total = 0
    it, last = istart, istart + n.niter
    while it <= last:
        stop = min(last, it + (-it) % n.freqbckp)
        hist = steps_with_energy(n, stop - it + 1)
        for j in range(stop - it + 1):
            EΩ, EΔ, Eβ, E, μ = (float(x) for x in hist[j])
            if verbose:
                print(f"iteration {it + j}")
                if n.f.g.ndims == 2:
                    print(f"Chemical potential : {μ}")
                print(f"Angular Momentum Energy : {EΩ}")
                print(f"Kinetic + Potential Energy : {EΔ}")
                print(f"Interaction Energy : {Eβ}")
                print(f"Total Energy: {E}")
            total += n.order
            res.append((n.order, total, EΩ, EΔ, Eβ, E))
        if stop % n.freqbckp == 0:
            save(stop)
        it = stop + 1
    return res


# ------------------------------------------------------ directional FFTs
def _fft(f: Field, out, src, axis: int, direction: int):
    _check_device_tensor(out, "out", f.ϕ.dtype, f.g.shape)
    _check_device_tensor(src, "src", f.ϕ.dtype, f.g.shape)
    c = f.ctx()
    L.check(c.lib.sfd_fft_axis(c.h, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                               axis, direction))


def mul_x(out, f: Field, src): _fft(f, out, src, 0, L.SFD_FORWARD)
def mul_y(out, f: Field, src): _fft(f, out, src, 1, L.SFD_FORWARD)
def mul_z(out, f: Field, src): _fft(f, out, src, 2, L.SFD_FORWARD)
def ldiv_x(out, f: Field, src): _fft(f, out, src, 0, L.SFD_BACKWARD)
def ldiv_y(out, f: Field, src): _fft(f, out, src, 1, L.SFD_BACKWARD)
def ldiv_z(out, f: Field, src): _fft(f, out, src, 2, L.SFD_BACKWARD)


def grid_x(f: Field):
    """Coordinates and wavenumbers broadcast in pen_x order (Plans.jl:1282-1320);
    all directional transforms here keep the pen_x layout."""
    co = f.coords()
    nd = f.g.ndims
    xis = []
    for a in range(nd):
        shp = [1] * nd
        shp[nd - 1 - a] = f.g.n[a]
        xis.append(fftfreq(f.g.n[a], f.g.Δ[a]).reshape(shp))
    return co + tuple(xis)


grid_y = grid_x
grid_z = grid_x


class Plan:
    """Plan(f; t=FFTPlan()) (Plans.jl:821-854, 958-999) on a HIP field: the
    scratch buffers of the FFT plan (ϕx_hat, a_tmpx, a_tmp2x and the y/z ones,
    Plans.jl:1078-1100), all in the pen_x layout, with mul_d / ldiv_d through
    sfd_fft_axis.  The plug of the reference's other FFT-based models."""

    def __init__(self, f: Field):
        self.f = f
        for d in "xyz"[:f.g.ndims]:
            for name in ("ϕ{}_hat", "a_tmp{}", "a_tmp2{}"):
                setattr(self, name.format(d), torch.empty_like(f.ϕ))
        rdt = torch.float32 if f.g.FT is np.float32 else torch.float64
        host = grid_x(f)
        self.grid = tuple(torch.as_tensor(np.ascontiguousarray(a), dtype=rdt, device=f.ϕ.device) for a in host)

    def mul(self, out, src, axis: int):
        _fft(self.f, out, src, axis, L.SFD_FORWARD)

    def ldiv(self, out, src, axis: int):
        _fft(self.f, out, src, axis, L.SFD_BACKWARD)


class GradientField:
    """GradientField(f; rotation=true) (Fields.jl): dx, dy[, dz], ddx, ddy[, ddz],
    rx, ry on the device."""

    def __init__(self, f: Field):
        for name in ("dx", "dy", "dz", "ddx", "ddy", "ddz", "rx", "ry")[:]:
            if name[-1] == "z" and f.g.ndims < 3:
                continue
            setattr(self, name, torch.empty_like(f.ϕ))


def computeDerivatives(gf: GradientField, p: Plan, ϕt) -> None:
    """computeDerivatives!(gf::GradientRot*, p, ϕt) (derivatives.jl:27-54, 90-126):
    per axis d, ϕd_hat = mul_d(ϕt); dd = ldiv_d(i ξd ϕd_hat), ddd = ldiv_d(-ξd² ϕd_hat);
    rx = ldiv_x(i y ξx ϕx_hat), ry = ldiv_y(-i x ξy ϕy_hat)."""
    nd = p.f.g.ndims
    co, xi = p.grid[:nd], p.grid[nd:]
    for a, d in enumerate("xyz"[:nd]):
        hat, tmp = getattr(p, f"ϕ{d}_hat"), getattr(p, f"a_tmp2{d}")
        p.mul(hat, ϕt, a)
        torch.mul(hat, 1j * xi[a], out=tmp)
        p.ldiv(getattr(gf, f"d{d}"), tmp, a)
        torch.mul(hat, -(xi[a] ** 2), out=tmp)
        p.ldiv(getattr(gf, f"dd{d}"), tmp, a)
        if a == 0 and nd >= 2:
            torch.mul(hat, 1j * co[1] * xi[0], out=tmp)
            p.ldiv(gf.rx, tmp, 0)
        if a == 1:
            torch.mul(hat, -1j * co[0] * xi[1], out=tmp)
            p.ldiv(gf.ry, tmp, 1)


# --------------------------------------------------- batched ground states
class BatchImagSolver:
    """Many independent 2D and 3D imaginary-time ground-state solves advanced
    together (sfd_batch_imag_solve): fields of equal shape share every sweep
    launch, each field keeps its own β, Ω, V, FP64 renormalisation factor,
    step count and convergence flag, and chunks of steps replay as hipGraphs.

    ``fields`` is a list of 2D or 3D Fields of one precision; ``params`` a
    list of GrossPitaevskiiParameters (one coeffΔ for all; Ω per field)."""

    def __init__(self, fields, params, Δt):
        _require_gpu()
        if len(fields) != len(params) or not fields:
            raise L.ArgError(L.SFD_ERR_ARG, "one parameter set per field")
        g0 = fields[0].g
        for f in fields:
            if f.g.ndims not in (2, 3):
                raise L.ArgError(L.SFD_ERR_ARG, "batched solves are 2D or 3D")
            if (f.g.FT, f.g.compute) != (g0.FT, g0.compute):
                raise L.ArgError(L.SFD_ERR_ARG, "all fields of a batch share one precision")
        cds = {p.coeffΔ for p in params}
        if len(cds) != 1:
            raise L.ArgError(L.SFD_ERR_ARG, "batched solves need one coeffΔ")
        self.fields, self.params, self.Δt = list(fields), list(params), float(Δt)
        self.coeffΔ = cds.pop()
        self.levels = None
        lib = L.load()
        self.lib = lib
        ni = len(fields)
        nd = (ctypes.c_int * ni)(*[f.g.ndims for f in fields])
        n = (ctypes.c_int64 * (3 * ni))(*[v for f in fields for v in (list(f.g.n) + [1] * (3 - f.g.ndims))])
        box = (ctypes.c_double * (6 * ni))(*[v for f in fields for a in range(3)
                                             for v in ((f.g.min[a], f.g.max[a]) if a < f.g.ndims else (0.0, 1.0))])
        if g0.FT is np.float64:
            prec = L.SFD_F64
        else:
            prec = L.SFD_F32X if g0.compute is np.float64 else L.SFD_F32
        h = ctypes.c_void_p()
        L.check(lib.sfd_batch_create_nd(ctypes.byref(h), ni, nd, n, box, prec,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        self.h = h

    def autotune(self):
        """Pick the fastest batched-sweep variant per shape group and axis
        (sfd_batch_autotune; the fields are not touched).  Returns a list of
        {axis: [ms per launch of each variant]} per shape group."""
        ni = len(self.fields)
        times = (ctypes.c_float * (24 * ni))()
        L.check(self.lib.sfd_batch_autotune(self.h, times))
        out = []
        for g in range(ni):
            row = {a: [t for t in times[24 * g + 8 * a:24 * g + 8 * a + 8] if t >= 0] for a in range(3)}
            row = {a: v for a, v in row.items() if v}
            if row:
                out.append(row)
        return out

    def run(self, max_steps=20000, tol=1e-9, check_every=100, use_graph=True, escalate=False, fp64_out=None):
        """Returns (steps taken per field, last ||Δψ|| per field).  With
        ``escalate`` (FP32 fields) each field finishes in FP64 from its FP32
        floor; ``fp64_out`` (a list of complex128 device tensors or None per
        field) receives the final states in FP64.  ``self.levels`` holds 1
        for the fields that finished in FP64."""
        ni = len(self.fields)
        psi = (ctypes.c_void_p * ni)(*[f.ϕ.data_ptr() for f in self.fields])
        vps = [p.pot.device_ptr() if p.pot is not None else None for p in self.params]
        V = (ctypes.c_void_p * ni)(*[v if v else None for v in vps])
        beta = (ctypes.c_double * ni)(*[p.β for p in self.params])
        om = (ctypes.c_double * ni)(*[p.Ω for p in self.params])
        steps = (ctypes.c_int64 * ni)()
        dpsi = (ctypes.c_double * ni)()
        lev = (ctypes.c_int * ni)()
        out64 = None
        if fp64_out is not None:
            if len(fp64_out) != ni:
                raise L.ArgError(L.SFD_ERR_ARG, "fp64_out needs one entry per field")
            for t, f in zip(fp64_out, self.fields):
                if t is not None:
                    _check_device_tensor(t, "fp64_out", torch.complex128, f.g.shape)
            out64 = (ctypes.c_void_p * ni)(*[t.data_ptr() if t is not None else None for t in fp64_out])
        flags = (L.SFD_BATCH_GRAPH if use_graph else 0) | (L.SFD_BATCH_ESCALATE if escalate else 0)
        L.check(self.lib.sfd_batch_imag_solve(self.h, psi, V, beta, om, self.coeffΔ, self.Δt, int(max_steps),
                                              float(tol), int(check_every), flags, steps, dpsi, out64, lev))
        self.levels = np.array(lev[:], dtype=np.int32)
        return np.array(steps[:], dtype=np.int64), np.array(dpsi[:])

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.sfd_batch_destroy(h)
            except Exception:
                pass
            self.h = None


# ------------------------------------------------------------- adjoint
def inner(f: Field, a, b) -> complex:
    """<a|b> = Σ conj(a) b ΠΔ (FP64) for tensors on f's grid."""
    _check_device_tensor(a, "a", f.ϕ.dtype, f.g.shape)
    _check_device_tensor(b, "b", f.ϕ.dtype, f.g.shape)
    c = f.ctx()
    out = (ctypes.c_double * 2)()
    L.check(c.lib.sfd_inner(c.h, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), out))
    return complex(out[0], out[1])


def _adjoint_call(n: _SplitBase, Vptr, beta: float, target, nsteps: int, checkpoint_every: int, want_state: bool):
    if n.mode != L.SFD_REAL_TIME or n.order != 2:
        raise L.ArgError(L.SFD_ERR_ARG, "the adjoint is defined for the real-time Strang split")
    f = n.f
    c = f.ctx()
    p = n.param
    tgt = target.ϕ if isinstance(target, Field) else target
    _check_device_tensor(tgt, "target", f.ϕ.dtype, f.g.shape)
    gV = torch.zeros(f.g.shape, dtype=torch.float64, device="cuda")
    gb, loss = ctypes.c_double(), ctypes.c_double()
    out = torch.empty_like(f.ϕ) if want_state else None
    L.check(c.lib.sfd_adjoint(c.h, ctypes.c_void_p(f.ϕ.data_ptr()), Vptr, ctypes.c_void_p(tgt.data_ptr()),
                              p.coeffΔ, float(beta), p.Ω, n.Δt, int(nsteps), int(checkpoint_every),
                              ctypes.c_void_p(gV.data_ptr()), ctypes.byref(gb), ctypes.byref(loss),
                              ctypes.c_void_p(out.data_ptr()) if want_state else None))
    return loss.value, gV, gb.value, out


class FidelityLoss(torch.autograd.Function):
    """L(V, β) = 1 - |<target|ψ(T)>|² after ``nsteps`` real-time Strang steps of
    model ``n`` from its current field, as a torch autograd function (the
    Python counterpart of a ChainRulesCore ``rrule``).  ``forward`` runs
    ``sfd_adjoint``, which forms dL/dV and dL/dβ along with the loss (√N
    checkpoint/recompute), and ``backward`` scales them by the incoming
    gradient.  ``V`` is a real tensor of the grid's shape, used in the field's
    precision; ``beta`` a 0-d tensor.  The model's own potential and β are not
    used."""

    @staticmethod
    def forward(ctx, V, beta, n, target, nsteps, checkpoint_every=0):
        f = n.f
        rdt = torch.float32 if f.g.FT is np.float32 else torch.float64
        if tuple(V.shape) != tuple(f.g.shape):
            raise L.ShapeError(L.SFD_ERR_SHAPE, f"V has shape {tuple(V.shape)}, the grid {tuple(f.g.shape)}")
        Vd = V.detach().to(device="cuda", dtype=rdt).contiguous()
        loss, gV, gb, _ = _adjoint_call(n, ctypes.c_void_p(Vd.data_ptr()), float(beta), target, nsteps,
                                        checkpoint_every, False)
        ctx.save_for_backward(gV, torch.tensor(gb, dtype=torch.float64, device=gV.device))
        ctx.dtypes = (V.dtype, beta.dtype if torch.is_tensor(beta) else torch.float64)
        return torch.tensor(loss, dtype=torch.float64, device=gV.device)

    @staticmethod
    def backward(ctx, g):
        gV, gb = ctx.saved_tensors
        dV, db = ctx.dtypes
        return (g * gV).to(dV), (g * gb).to(db), None, None, None, None


def fidelity_loss(n: _SplitBase, V, beta, target, nsteps: int, checkpoint_every: int = 0):
    """``FidelityLoss.apply``: a differentiable L(V, β) for torch autograd."""
    if not torch.is_tensor(beta):
        beta = torch.tensor(float(beta), dtype=torch.float64, device="cuda")
    return FidelityLoss.apply(V, beta, n, target, nsteps, checkpoint_every)


def adjoint_gradient(n: _SplitBase, target, nsteps: int, checkpoint_every: int = 0, want_state: bool = False):
    """Reverse-mode gradient of L = 1 - |<target|ψ(T)>|² through ``nsteps``
    real-time Strang steps of model ``n`` from its current field (which is
    left unchanged).  Returns (L, dL/dV as an FP64 device tensor, dL/dβ[, ψ(T)]).
    Forward states are recomputed from √N checkpoints (or every
    ``checkpoint_every`` steps)."""
    loss, gV, gb, out = _adjoint_call(n, _vptr(n.param), n.param.β, target, nsteps, checkpoint_every, want_state)
    res = (loss, gV, gb)
    return res + (out,) if want_state else res


def project(f: Field, field, basis) -> np.ndarray:
    """out[j] = Σ_x field[x]·basis[j, x] (FP64 device tensors; basis shape
    (nb,) + grid shape): the chain rule from dL/dV to V = Σ_j a_j basis_j."""
    _check_device_tensor(field, "field", torch.float64, f.g.shape)
    if not torch.is_tensor(basis) or basis.dim() != len(f.g.shape) + 1:
        raise L.ShapeError(L.SFD_ERR_SHAPE, f"basis must have shape (nb,) + {tuple(f.g.shape)}")
    nb = basis.shape[0]
    _check_device_tensor(basis, "basis", torch.float64, (nb,) + tuple(f.g.shape))
    c = f.ctx()
    out = (ctypes.c_double * nb)()
    L.check(c.lib.sfd_project(c.h, ctypes.c_void_p(field.data_ptr()), ctypes.c_void_p(basis.data_ptr()), nb, out))
    return np.array(out[:])
