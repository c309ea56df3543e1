"""Host-side mirror of SPPMManager (src/main/java/com/retrocam/scene/SPPMManager.java)
over the C ABI: tracePhotons / gatherRadiance / updateRadius / reset, plus a
readback of the photon map for tests."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

HASH_SIZE = 1 << 22        # PhotonMap.HASH_SIZE (PhotonMap.java:41)
MAX_PHOTONS = 2_000_000    # PhotonMap.MAX_PHOTONS (PhotonMap.java:44)


class SPPMManager:
    def __init__(self, ctx, init_radius: float = 0.05, fp64_radius: bool = False):
        self.ctx = ctx
        self._lib = ctx._lib
        self.photon_count = 0
        self.hash_size = HASH_SIZE
        ctx.check(self._lib.rc_sppm_set_precision(ctx.h, 1 if fp64_radius else 0))
        self.reset(init_radius)

    def reset(self, init_radius: float):
        self.ctx.check(self._lib.rc_sppm_reset(self.ctx.h, float(init_radius)))

    def trace_photons(self, photons_per_iter: int, hash_size: int = HASH_SIZE, cap: int = MAX_PHOTONS):
        # SPPMManager.java:83: photonCount = min(photonsPerIter, MAX_PHOTONS)
        self.photon_count = min(int(photons_per_iter), int(cap))
        self.hash_size = int(hash_size)
        self.ctx.check(self._lib.rc_sppm_trace(self.ctx.h, self.photon_count, self.hash_size))

    def gather_radiance(self, camera):
        cam = camera.to_c()
        self.ctx.check(self._lib.rc_sppm_gather(self.ctx.h, C.byref(cam)))

    def update_radius(self, alpha: float):
        self.ctx.check(self._lib.rc_sppm_update_radius(self.ctx.h, float(alpha)))

    def state(self):
        r = C.c_float()
        it = C.c_int32()
        r64 = C.c_double()
        self.ctx.check(self._lib.rc_sppm_get_state(self.ctx.h, C.byref(r), C.byref(it), C.byref(r64)))
        return float(r.value), int(it.value), float(r64.value)

    @property
    def search_radius(self) -> float:
        return self.state()[0]

    @property
    def iteration(self) -> int:
        return self.state()[1]

    def stored_photons(self) -> int:
        """Photons the last trace stored (the insertion counter, capped by the map size)."""
        cnt = np.zeros(1, np.uint32)
        self.ctx.check(self._lib.rc_sppm_read(self.ctx.h, None, None, None, cnt.ctypes.data_as(C.POINTER(C.c_uint32))))
        return min(int(cnt[0]), self.photon_count)

    def read_map(self):
        P, Hs = self.photon_count, self.hash_size
        photons = np.zeros((P, 12), np.float32)
        heads = np.zeros(Hs, np.int32)
        nxt = np.zeros(P, np.int32)
        cnt = np.zeros(1, np.uint32)
        self.ctx.check(self._lib.rc_sppm_read(self.ctx.h, N.fptr(photons), heads.ctypes.data_as(C.POINTER(C.c_int32)),
                                              nxt.ctypes.data_as(C.POINTER(C.c_int32)),
                                              cnt.ctypes.data_as(C.POINTER(C.c_uint32))))
        return photons, heads, nxt, int(cnt[0])
