"""AthenaK meshblock snapshots, resident in HBM (reference: grmhd/athenak.py).

``AthenakFluidModel`` keeps the reference's constructor, attributes and
sampler methods.  The ghost-padded meshblock layout is built on the GPU
(``mk_snapshot_create``) and stays resident; the sampler methods evaluate
block lookup, trilinear interpolation, four-velocity, magnetic four-vector
and scalars in a HIP kernel instead of the reference's host mask loop.
``from_datasets`` builds a model from in-memory AthenaK datasets (as read
from an .athdf file), which is how the synthetic snapshots are loaded.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from mahakala_amd import _lib
from mahakala_amd._array import wrap
from mahakala_amd.geodesics import metric, imetric
from mahakala_amd.grmhd.grmhd import GRMHDFluidModel

_CANON = ("dens", "eint", "velx", "vely", "velz", "bcc1", "bcc2", "bcc3")


def vec_metric(X, bhspin):
    """metric over points (n, 4) -> (n, 4, 4)  (athenak.py:38-40)."""
    return metric(X, bhspin)


def vec_imetric(X, bhspin):
    """inverse metric over points (n, 4) -> (n, 4, 4)  (athenak.py:43-45)."""
    return imetric(X, bhspin)


def read_athdf(filename):
    """Datasets of an AthenaK .athdf file as a dict (athenak.py:79-103).

    Uses h5py when it is importable, else the native reader
    (libmk_athdf.so, the HDF5 C library); the two give the same arrays."""
    try:
        import h5py
    except ImportError:  # the image's main interpreter has no h5py
        from mahakala_amd import _athdf
        try:
            return _athdf.read_athdf_native(filename)
        except _athdf.AthdfError as exc:
            if not os.path.exists(_athdf.LIB_PATH):
                raise ImportError("reading .athdf files needs h5py or libmk_athdf.so; build the model "
                                  "with AthenakFluidModel.from_datasets instead") from exc
            raise
    with h5py.File(filename, "r") as f:
        ds = {k: np.array(f[k]) for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "uov", "B",
                                          "LogicalLocations", "Levels")}
        ds["VariableNames"] = [n.decode("utf-8") for n in f.attrs["VariableNames"]]
    return ds


class _DeviceSnapshot:
    """Owner of an mk_snapshot handle."""

    def __init__(self, prims, var_index, levels, logical, x1f, x2f, x3f, x1v, x2v, x3v):
        nprim, nmb, n, _, _ = prims.shape
        self.nmb, self.n = nmb, n
        h = _lib.vp()
        arr = lambda a, t: np.ascontiguousarray(a, dtype=t)  # noqa: E731
        keep = [arr(prims, np.float32), arr(var_index, np.int32), arr(levels, np.int32),
                arr(logical, np.int32)] + [arr(v, np.float64) for v in (x1f, x2f, x3f, x1v, x2v, x3v)]
        _lib.check(_lib.lib().mk_snapshot_create(keep[0].ctypes.data, keep[1].ctypes.data, nmb, n,
                                                 keep[2].ctypes.data, keep[3].ctypes.data,
                                                 *[k.ctypes.data for k in keep[4:]], C.byref(h)))
        self.handle = h.value
        nbytes = C.c_size_t()
        _lib.check(_lib.lib().mk_snapshot_bytes(self.handle, C.byref(nbytes)))
        self.nbytes = nbytes.value

    def download(self):
        n2 = self.n + 2
        out = np.empty((self.nmb, n2, n2, n2, 8), dtype=np.float32)
        _lib.check(_lib.lib().mk_snapshot_download(self.handle, out.ctypes.data))
        return out

    def __del__(self):
        if getattr(self, "handle", None) and _lib._lib is not None:
            _lib._lib.mk_snapshot_free(self.handle)
            self.handle = None


class AthenakFluidModel(GRMHDFluidModel):
    """AthenaK GRMHD snapshot sampled along geodesics (athenak.py:48-812)."""

    def __init__(self, grmhd_filename, bhspin, fluid_gamma=None):
        self.load_athenak_meshblocks(grmhd_filename)
        self.bhspin = bhspin
        self.fluid_gamma = fluid_gamma

    @classmethod
    def from_datasets(cls, datasets, bhspin, fluid_gamma=None):
        """Model from in-memory .athdf datasets: uov (5, nmb, n, n, n), B (3, ...),
        x{1,2,3}{v,f}, LogicalLocations, Levels, VariableNames."""
        obj = cls.__new__(cls)
        obj._load_datasets(datasets)
        obj.bhspin = bhspin
        obj.fluid_gamma = fluid_gamma
        return obj

    def get_index_for_primitive_by_name(self, prim):
        """Index of a primitive in the file's variable list, -1 if absent (athenak.py:55-65)."""
        prim = prim.lower().strip()
        names = [v.lower().strip() for v in self.variable_names]
        if prim in names:
            return names.index(prim)
        return -1

    def load_athenak_meshblocks(self, grmhd_filename):
        """Load an .athdf file and build the resident snapshot (athenak.py:67-158)."""
        self._load_datasets(read_athdf(grmhd_filename))

    def _load_datasets(self, ds):
        uov = np.asarray(ds["uov"])
        B = np.asarray(ds["B"])
        self.variable_names = np.array(list(ds["VariableNames"]))
        nprim, nmb, nk, nj, ni = uov.shape
        if not (nk == nj == ni):
            raise ValueError("meshblocks must be cubic")
        self.nprim_all = 8
        prims = np.concatenate([uov, B], axis=0).astype(np.float32, copy=False)
        if prims.shape[0] != self.nprim_all:
            raise ValueError("expected 8 primitives (uov + B)")
        # rows of the canonical primitives; a missing name reads row -1 (the
        # last one), as indexing with the reference's -1 does
        self._var_index = np.array([self.get_index_for_primitive_by_name(nm) % self.nprim_all for nm in _CANON],
                                   dtype=np.int32)
        self.Levels = np.asarray(ds["Levels"])
        self.LogicalLocations = np.asarray(ds["LogicalLocations"])
        self.x1v, self.x2v, self.x3v = (np.asarray(ds[k]) for k in ("x1v", "x2v", "x3v"))
        self.x1f, self.x2f, self.x3f = (np.asarray(ds[k]) for k in ("x1f", "x2f", "x3f"))
        self.mb_index_map = {}
        for mb in range(nmb):
            key = (int(self.Levels[mb]),) + tuple(int(v) for v in self.LogicalLocations[mb])
            self.mb_index_map[key] = mb
        self._snap = _DeviceSnapshot(prims, self._var_index, self.Levels, self.LogicalLocations, self.x1f,
                                     self.x2f, self.x3f, self.x1v, self.x2v, self.x3v)

    @property
    def device_snapshot(self):
        """The resident snapshot (used by the fused image kernel)."""
        return self._snap

    @property
    def all_meshblocks(self):
        """Ghost-padded blocks (nmb, 8, n+2, n+2, n+2), rows in file variable order."""
        cells = self._snap.download()  # (nmb, k, j, i, canonical)
        out = np.zeros((cells.shape[0], self.nprim_all) + cells.shape[1:4])
        for c, row in enumerate(self._var_index):
            out[:, row] = cells[..., c]
        return out

    def _check_indices(self):
        iU1, iU2, iU3 = (self.get_index_for_primitive_by_name(n) for n in ("velx", "vely", "velz"))
        iB1, iB2, iB3 = (self.get_index_for_primitive_by_name(n) for n in ("bcc1", "bcc2", "bcc3"))
        if iU2 != iU1 + 1 or iU3 != iU1 + 2:
            raise ValueError("Velocity indices are not as expected")
        if iB2 != iB1 + 1 or iB3 != iB1 + 2:
            raise ValueError("Magnetic field indices are not as expected")

    def _states(self, S):
        S = np.ascontiguousarray(np.asarray(S, dtype=np.float64))
        if S.ndim != 3 or S.shape[-1] != 8:
            raise ValueError("S must have shape (nsteps, npx, 8)")
        return S

    def get_prims_from_geodesics(self, S, profile=False):
        """Primitives at the points of S (nsteps, npx, 8), zero outside the
        domain: dict dens, u, U1, U2, U3, B1, B2, B3  (athenak.py:527-637)."""
        S = self._states(S)
        nsteps, npx, _ = S.shape
        n = nsteps * npx
        t0 = time.time()
        d_S = _lib.DeviceBuffer.from_array(S)
        out = _lib.DeviceBuffer(8 * 8 * n)
        _lib.check(_lib.lib().mk_sample_prims(self._snap.handle, d_S.ptr, n, out.ptr, None))
        P = out.to_array((8, nsteps, npx))
        t1 = time.time()
        if profile:
            print(f"Time to compute meshblock indices: {0.0}")
            print(f"Time to compute primitives: {t1 - t0}")
        keys = ("dens", "u", "U1", "U2", "U3", "B1", "B2", "B3")
        return {k: wrap(P[c]) for c, k in enumerate(keys)}

    def get_fluid_scalars_from_geodesics(self, S, fallback_pitch_angle=np.pi / 3., profile=False):
        """dens, u, pitch_angle, kdotu, b at the points of S  (athenak.py:639-812).

        One GPU thread per (step, ray): leaf lookup on the level grid,
        trilinear interpolation of the ghost-padded cells, normal-observer
        four-velocity, magnetic four-vector b, k.u, k.b, b.b, pitch angle.
        """
        self._check_indices()
        S = self._states(S)
        nsteps, npx, _ = S.shape
        n = nsteps * npx
        t0 = time.time()
        d_S = _lib.DeviceBuffer.from_array(S)
        out = _lib.DeviceBuffer(8 * 5 * n)
        _lib.check(_lib.lib().mk_sample_scalars(self._snap.handle, d_S.ptr, n, float(self.bhspin),
                                                float(fallback_pitch_angle), out.ptr, None))
        sc = out.to_array((5, nsteps, npx))
        t1 = time.time()
        if profile:
            print(f"Time to compute meshblock indices: {0.0}")
            print(f"Time to compute scalar data: {t1 - t0}")
        return dict(dens=wrap(sc[0]), u=wrap(sc[1]), pitch_angle=wrap(sc[2]), kdotu=wrap(sc[3]), b=wrap(sc[4]))

    def block_index(self, S):
        """Leaf id of each point of S (..., >=4), -1 outside (athenak.py:657-670)."""
        S = np.ascontiguousarray(np.asarray(S, dtype=np.float64))
        lead = S.shape[:-1]
        n = int(np.prod(lead))
        d_S = _lib.DeviceBuffer.from_array(S)
        out = _lib.DeviceBuffer(4 * n)
        _lib.check(_lib.lib().mk_block_index(self._snap.handle, d_S.ptr, n, S.shape[-1], out.ptr, None))
        return out.to_array(lead, np.int32)
