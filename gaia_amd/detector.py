"""Host interface of the GPU contact detector, mirroring GAIA::ClothContactDetector
(Simulator/Modules/CollisionDetector/ClothContactDetector.h:133-182): initialize(meshes) ->
updateBVH(REFIT / LOW) -> contactQueryVF / contactQueryEE over every vertex and edge, face
minimum distances, and the conservative bounds of applyCollisionDetection. The work runs in the
LBVH kernels of libgaia_vbd.so."""
import ctypes

import numpy as np

from ._lib import f32, fptr, i32, iptr, lib

VF_DTYPE = np.dtype([("face", np.int32), ("type", np.int32), ("prim", np.int32), ("d", np.float32)])
EE_DTYPE = np.dtype([("edge", np.int32), ("mu1", np.float32), ("mu2", np.float32), ("d", np.float32)])


class ContactDetector:
    """meshes: sequence of (positions (n, 3), faces (f, 3)); ids are global over the meshes."""

    def __init__(self, ctx, meshes, max_query_dis=1.2, vf_capacity=8, ee_capacity=8):
        self.ctx = ctx
        h = ctypes.c_int32()
        ctx.check(lib().gvbd_cd_create(ctx.handle, ctypes.c_float(max_query_dis), vf_capacity, ee_capacity,
                                       ctypes.byref(h)))
        self.cd = h.value
        self.max_query_dis = np.float32(max_query_dis)
        self.vf_capacity, self.ee_capacity = vf_capacity, ee_capacity
        for pos, faces in meshes:
            p = f32(pos).reshape(-1)
            f = i32(faces).reshape(-1)
            m = ctypes.c_int32()
            ctx.check(lib().gvbd_cd_add_mesh(ctx.handle, self.cd, len(p) // 3, fptr(p), len(f) // 3, iptr(f),
                                             ctypes.byref(m)))
        ctx.check(lib().gvbd_cd_finalize(ctx.handle, self.cd))
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ctx.check(lib().gvbd_cd_sizes(ctx.handle, self.cd, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        self.nv, self.nf, self.ne = a.value, b.value, c.value
        self.num_meshes = len(meshes)

    def mesh_offsets(self, mesh):
        v, f, e = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(lib().gvbd_cd_mesh_offsets(self.ctx.handle, self.cd, mesh, ctypes.byref(v), ctypes.byref(f),
                                                  ctypes.byref(e)))
        return v.value, f.value, e.value

    def topology(self):
        edges = np.zeros((self.ne, 4), np.int32)
        fe = np.zeros((self.nf, 3), np.int32)
        self.ctx.check(lib().gvbd_cd_topology(self.ctx.handle, self.cd, iptr(edges), iptr(fe)))
        return edges, fe

    def set_positions(self, mesh, positions):
        p = f32(positions).reshape(-1)
        self.ctx.check(lib().gvbd_cd_set_positions(self.ctx.handle, self.cd, mesh, fptr(p)))

    def positions(self):
        out = []
        for m in range(self.num_meshes):
            v0 = self.mesh_offsets(m)[0]
            v1 = self.mesh_offsets(m + 1)[0] if m + 1 < self.num_meshes else self.nv
            p = np.zeros(3 * (v1 - v0), np.float32)
            self.ctx.check(lib().gvbd_cd_get_positions(self.ctx.handle, self.cd, m, fptr(p)))
            out.append(p.reshape(-1, 3))
        return np.concatenate(out)

    def perturb(self, seed, step, sigma, drift=None):
        d = None if drift is None else f32(drift).reshape(-1)
        self.ctx.check(lib().gvbd_cd_perturb(self.ctx.handle, self.cd, seed, step, ctypes.c_float(sigma),
                                             fptr(d) if d is not None else None))

    def updateBVH(self, rebuild=False):
        self.ctx.check(lib().gvbd_cd_update_bvh(self.ctx.handle, self.cd, 1 if rebuild else 0))

    def query(self, record=False):
        self.ctx.check(lib().gvbd_cd_query(self.ctx.handle, self.cd, 1 if record else 0))

    def counts(self):
        c = np.zeros(5, np.int64)
        self.ctx.check(lib().gvbd_cd_counts(self.ctx.handle, self.cd, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(vf_pairs=int(c[0]), vf_contacts=int(c[1]), ee_contacts=int(c[2]), vf_overflow=int(c[3]),
                    ee_overflow=int(c[4]))

    def recorded(self):
        """Uncapped lists of the last query(record=True), in the oracle's layout."""
        c = self.counts()
        L, h, cd = lib(), self.ctx.handle, self.cd
        vf = np.zeros((c["vf_pairs"], 2), np.int32)
        vfd = np.zeros(c["vf_pairs"], np.float32)
        self.ctx.check(L.gvbd_cd_vf_pairs(h, cd, iptr(vf), fptr(vfd)))
        rec = np.zeros((c["vf_contacts"], 4), np.int32)
        recd = np.zeros(c["vf_contacts"], np.float32)
        self.ctx.check(L.gvbd_cd_vf_contacts(h, cd, iptr(rec), fptr(recd)))
        ee = np.zeros((c["ee_contacts"], 2), np.int32)
        mu = np.zeros((c["ee_contacts"], 2), np.float32)
        eed = np.zeros(c["ee_contacts"], np.float32)
        self.ctx.check(L.gvbd_cd_ee_contacts(h, cd, iptr(ee), fptr(mu), fptr(eed)))
        return dict(vf_pairs=vf, vf_pair_d=vfd, vf_contacts=rec, vf_contact_d=recd, ee_contacts=ee, ee_mu=mu,
                    ee_d=eed)

    def results(self):
        """Capped per-vertex VF and per-edge EE contacts of the last query."""
        vc = np.zeros(self.nv, np.int32)
        vr = np.zeros(self.nv * self.vf_capacity, VF_DTYPE)
        self.ctx.check(lib().gvbd_cd_vf_results(self.ctx.handle, self.cd, iptr(vc), vr.ctypes.data))
        ec = np.zeros(self.ne, np.int32)
        er = np.zeros(self.ne * self.ee_capacity, EE_DTYPE)
        self.ctx.check(lib().gvbd_cd_ee_results(self.ctx.handle, self.cd, iptr(ec), er.ctypes.data))
        return vc, vr.reshape(self.nv, self.vf_capacity), ec, er.reshape(self.ne, self.ee_capacity)

    def min_distances(self):
        a, b, c = np.zeros(self.nv, np.float32), np.zeros(self.nf, np.float32), np.zeros(self.ne, np.float32)
        self.ctx.check(lib().gvbd_cd_min_distances(self.ctx.handle, self.cd, fptr(a), fptr(b), fptr(c)))
        return a, b, c

    def bounds(self, thickness, relaxation):
        b = np.zeros(self.nv, np.float32)
        self.ctx.check(lib().gvbd_cd_bounds(self.ctx.handle, self.cd, ctypes.c_float(thickness),
                                            ctypes.c_float(relaxation), fptr(b)))
        return b

    def destroy(self):
        if self.cd is not None and self.ctx._h:
            lib().gvbd_cd_destroy(self.ctx.handle, self.cd)
        self.cd = None
