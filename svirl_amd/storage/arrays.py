"""Host/device mirrored arrays with the flatten convention of
svirl/storage/arrays.py:327-365: host arrays are shaped (Nx, Ny) and indexed
[i, j]; device storage is flat with i fastest (``reshape(arr.T, n)``).
Vector storage ``shape=[(Nxa, Nya), (Nxb, Nyb)]`` packs [a; b]."""
from warnings import warn

import numpy as np

from svirl_amd._native import DeviceArray


class GArray:
    """Mirrored host/device array (the interface of svirl's GArray,
    svirl/storage/arrays.py:11-609, on a libglx device buffer)."""

    on_device = 'd'
    on_host_and_device = 'hd'

    def __init__(self, nelements=None, dtype=np.float64, shape=None, on=None, like=None,
                 like_init=True, name='array-0', suppress_warnings=False):
        if on is not None and on not in (self.on_device, self.on_host_and_device):
            raise ValueError("`on` must be on_device or on_host_and_device")
        self.name = str(name)
        self._data = None
        self._gdata = None
        self._data_v = None
        self._nelements_v = None
        self._sync_status = 0
        self._ndim = 1
        if shape is not None and isinstance(shape, list):
            self._ndim = len(shape)
        if like is not None:
            if isinstance(like, GArray):
                src = like.get_d_obj().get()
                self._shape = like.shape
                self._dtype = np.dtype(like.dtype)
            elif isinstance(like, DeviceArray):
                src = like.get()
                self._shape = (like.size, 1)
                self._dtype = like.dtype
            else:
                arr = np.asarray(like)
                self._shape = arr.shape
                self._dtype = arr.dtype
                src = self._flatten(arr, arr.size) if like_init else None
            self._nelements = int(np.prod(self._shape))
            self._on = on or self.on_host_and_device
            self._gdata = DeviceArray(self._nelements, self._dtype)
            if like_init and src is not None:
                src = np.asarray(src, dtype=self._dtype).ravel()
                self._gdata.set(src)
                self._data = src.copy() if self._on == self.on_host_and_device else None
            else:
                self._data = np.zeros(self._nelements, dtype=self._dtype) if self._on == self.on_host_and_device else None
        else:
            if nelements is None and shape is None:
                raise ValueError("specify nelements or shape")
            self._dtype = np.dtype(dtype)
            if np.dtype(dtype) not in (np.float32, np.float64, np.complex64, np.complex128):
                raise TypeError("dtype must be float32/64 or complex64/128")
            self._on = on or self.on_host_and_device
            if shape is not None:
                self._shape = shape
                if self._ndim == 1:
                    self._nelements = int(np.prod(shape))
                else:
                    sizes = [int(np.prod(s)) for s in shape]
                    self._nelements = int(np.sum(sizes))
                    self._nelements_v = np.cumsum(sizes)
            else:
                self._shape = (int(nelements), 1)
                self._nelements = int(nelements)
            self._gdata = DeviceArray(self._nelements, self._dtype)
            if self._on == self.on_host_and_device:
                self._data = np.zeros(self._nelements, dtype=self._dtype)
                self._gdata.set(self._data)
            else:
                self._gdata.fill(0)
        if self._ndim > 1 and self._data is not None:
            self._data_v = np.split(self._data, self._nelements_v[:-1])

    # -- properties
    @property
    def size(self):
        return self._nelements

    @property
    def dtype(self):
        return self._dtype

    @property
    def shape(self):
        return self._shape

    @property
    def on(self):
        return self._on

    # -- sync status (-1: host newer, 0: synced, 1: device newer)
    def sync(self):
        if self._on != self.on_host_and_device:
            return
        if self._sync_status > 0:
            self._data[:] = self._gdata.get()
        elif self._sync_status < 0:
            self._gdata.set(self._data)
        self._sync_status = 0

    def sync_d(self):
        """Make the device copy current (host-to-device copy if the host is
        newer); a newer device copy stays where it is. For kernels that only
        read the device copy."""
        if self._on == self.on_host_and_device and self._sync_status < 0:
            self._gdata.set(self._data)
            self._sync_status = 0

    def need_dtoh_sync(self):
        if self._on == self.on_host_and_device:
            self._sync_status = 1

    def need_htod_sync(self):
        if self._on == self.on_host_and_device:
            self._sync_status = -1

    def synced(self):
        self._sync_status = 0

    # -- layout helpers
    @staticmethod
    def _flatten(arr, size):
        return np.reshape(np.asarray(arr).T, size)

    @staticmethod
    def _unflatten(arr, shape):
        return np.reshape(arr, tuple(reversed(shape))).T

    # -- getters
    def get_h(self, sync=True):
        if sync:
            self.sync()
        if self._ndim > 1:
            return None
        return self._unflatten(self._data, self._shape)

    def get_d(self, sync=True):
        if sync:
            self.sync()
        if self._ndim > 1:
            return None
        return self._unflatten(self._gdata.get(), self._shape)

    def get_d_obj(self, sync=False):
        if sync:
            self.sync()
        return self._gdata

    def device_pool(self):
        """Three same-sized device buffers, one of them the current device
        copy, and the index of that one. The TD solver rotates the field
        among them and the CG writes its updated state into one that is not
        current, so moving between solvers allocates nothing and never
        writes the buffer a kernel reads. A device copy from elsewhere
        starts a new set."""
        cur = self._gdata
        pool = getattr(self, '_pool', None)
        if pool is None or all(b is not cur for b in pool):
            pool = [cur, DeviceArray(cur.size, cur.dtype), DeviceArray(cur.size, cur.dtype)]
            self._pool = pool
        return pool, [k for k, b in enumerate(pool) if b is cur][0]

    def get_vec_h(self, sync=True):
        if sync:
            self.sync()
        if self._ndim > 1:
            return (self._unflatten(self._data_v[0], self._shape[0]),
                    self._unflatten(self._data_v[1], self._shape[1]))
        return (None, None)

    # -- setters
    def set_h(self, arr):
        arr = np.asarray(arr)
        if arr.size != self._nelements:
            warn("Size mismatch: Storage has size %d but input array has: %d" % (self._nelements, arr.size))
            return
        np.copyto(self._data, self._flatten(arr, self._nelements))
        self.need_htod_sync()

    def set_vec_h(self, arr_a, arr_b):
        assert arr_a.shape == self._shape[0]
        assert arr_b.shape == self._shape[1]
        if self._data is not None:
            np.copyto(self._data_v[0], self._flatten(arr_a, arr_a.size))
            np.copyto(self._data_v[1], self._flatten(arr_b, arr_b.size))
            self.need_htod_sync()

    def free(self):
        if self._gdata is not None:
            self._gdata.free()
            self._gdata = None
        self._data = None
        self._data_v = None
        self._nelements = 0
        self._shape = None
