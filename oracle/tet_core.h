// Shared by the tet and cloth oracles: rest-pose precompute, vertex -> tet incidence and the
// per-vertex neo-Hookean force/Hessian of the reference's CPU tet path.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked by the product.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace tetcore {

// Velocity damping of one vertex after v = (x - x_prev) / dt: dampVelocities
// (VBD/VBDPhysics.cpp:2529-2555, tet meshes, zone coordinate = the position's y) and
// updateVelocity (VBDCloth/VBDClothPhysics.cpp:2325-2351, cloth meshes, zone coordinate = the
// velocity's y). In the no-gravity zone (hasNoGravZone, maxVelocityMagnitude > 0 and the zone
// coordinate above noGravZoneThreshold) the speed is only capped at maxVelocityMagnitude;
// elsewhere |v| -> max(|v| expDamp - constDamp, 0) with speeds of at most 1e-6 zeroed.
template <class R>
inline void damp_velocity(R* v, bool noGravZone, R zoneCoord, R threshold, R maxVel, R expDamp, R constDamp) {
    const R vMag = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (noGravZone && maxVel > R(0) && zoneCoord > threshold) {
        if (vMag > maxVel) {
            const R s = maxVel / vMag;
            for (int d = 0; d < 3; ++d) v[d] *= s;
        }
    } else if (vMag > R(1e-6)) {
        R vNew = vMag * expDamp - constDamp;
        vNew = vNew > R(1e-6) ? vNew : R(0);
        for (int d = 0; d < 3; ++d) v[d] *= vNew / vMag;
    } else {
        for (int d = 0; d < 3; ++d) v[d] *= R(0);
    }
}

// ---------------------------------------------------------------------------
// 3x3 helpers, column-major like the reference's Eigen Mat3 (Types.h:76-127).
// m[r + 3*c]
// ---------------------------------------------------------------------------
template <class R> inline R det3(const R* m) {
    // cofactor expansion along the first column
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
           m[2] * (m[3] * m[7] - m[6] * m[4]);
}

template <class R> inline bool inv3(const R* m, R* out) {
    R c00 = m[4] * m[8] - m[7] * m[5];
    R c01 = -(m[1] * m[8] - m[7] * m[2]);
    R c02 = m[1] * m[5] - m[4] * m[2];
    R det = m[0] * c00 + m[3] * c01 + m[6] * c02;
    if (det == R(0)) return false;
    R inv = R(1) / det;
    // out = adj(m)/det, adj = transpose of cofactor matrix
    out[0] = c00 * inv;
    out[1] = c01 * inv;
    out[2] = c02 * inv;
    out[3] = -(m[3] * m[8] - m[6] * m[5]) * inv;
    out[4] = (m[0] * m[8] - m[6] * m[2]) * inv;
    out[5] = -(m[0] * m[5] - m[3] * m[2]) * inv;
    out[6] = (m[3] * m[7] - m[6] * m[4]) * inv;
    out[7] = -(m[0] * m[7] - m[6] * m[1]) * inv;
    out[8] = (m[0] * m[4] - m[3] * m[1]) * inv;
    return true;
}

// Own closed-form 3x3 solve used in place of CuMatrix::solve3x3_psd_stable
// (VBDPhysics.cpp:2449). CuMatrix is an empty submodule, so its rounding and
// failure threshold are unknown: parity unpinned. Definition used here and in
// the HIP kernels: adjugate / determinant; failure iff det is not finite or
// |det| <= 1e-12 * |h00 h11 h22|.

// TetMeshFEM.cpp:409-429: Dm = [x1-x0, x2-x0, x3-x0]; DmInv = Dm^-1; vol = det(Dm)/6;
// every tet vertex gets 0.25*vol*density of lumped mass.
template <class R>
int tet_prepare(int nv, const R* pos, int nt, const int32_t* tets, R density, R* dminv, R* vol, R* mass) {
    for (int i = 0; i < nv; ++i) mass[i] = R(0);
    for (int t = 0; t < nt; ++t) {
        const int32_t* tv = tets + 4 * t;
        R Dm[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Dm[r + 3 * c] = pos[3 * tv[c + 1] + r] - pos[3 * tv[0] + r];
        if (!inv3(Dm, dminv + 9 * t)) return -1;
        R v = det3(Dm) / R(6);
        vol[t] = v;
        // `vertexMass += 0.25 * vol * density` evaluates in double, then rounds
        for (int k = 0; k < 4; ++k) mass[tv[k]] = (R)((double)mass[tv[k]] + 0.25 * (double)v * (double)density);
    }
    return 0;
}

// VBD_NeoHookean.cpp:126-152 assembleVertexVForceAndHessian.
template <class R>
inline void assemble_vertex(const R* dEdF, const R* d2E, R m1, R m2, R m3, R* f, R* h) {
    f[0] = dEdF[0] * m1 + dEdF[3] * m2 + dEdF[6] * m3;
    f[1] = dEdF[1] * m1 + dEdF[4] * m2 + dEdF[7] * m3;
    f[2] = dEdF[2] * m1 + dEdF[5] * m2 + dEdF[8] * m3;
    R HL[3][9];
    for (int c = 0; c < 9; ++c) {
        for (int r = 0; r < 3; ++r)
            HL[r][c] = d2E[r + 9 * c] * m1 + d2E[(r + 3) + 9 * c] * m2 + d2E[(r + 6) + 9 * c] * m3;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) h[r + 3 * c] = HL[r][c] * m1 + HL[r][c + 3] * m2 + HL[r][c + 6] * m3;
}

// VBD_NeoHookean.cpp:379-581 accumlateMaterialForceAndHessian2, per incident tet:
// the full 9x9 d2E/dF2 of the stable neo-Hookean energy, contracted with the
// vertex's row m of the shape matrix, plus damping. Adds into force/hess.
template <class R>
void neohookean_vertex(const R* x4, const R* DmInv, R A, R miu, R lmbd, int order, R dampH, R dampD, R dt,
                       const R* disp, R* force, R* hess) {
    const R a = R(1) + miu / lmbd;
    R Ds[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Ds[r + 3 * c] = x4[3 * (c + 1) + r] - x4[r];
    R F[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            F[r + 3 * c] = Ds[r + 0] * DmInv[0 + 3 * c] + Ds[r + 3] * DmInv[1 + 3 * c] + Ds[r + 6] * DmInv[2 + 3 * c];
    const R detF = det3(F);
    const R F1_1 = F[0], F2_1 = F[1], F3_1 = F[2];
    const R F1_2 = F[3], F2_2 = F[4], F3_2 = F[5];
    const R F1_3 = F[6], F2_3 = F[7], F3_3 = F[8];
    R g[9] = {F2_2 * F3_3 - F2_3 * F3_2, F1_3 * F3_2 - F1_2 * F3_3, F1_2 * F2_3 - F1_3 * F2_2,
              F2_3 * F3_1 - F2_1 * F3_3, F1_1 * F3_3 - F1_3 * F3_1, F1_3 * F2_1 - F1_1 * F2_3,
              F2_1 * F3_2 - F2_2 * F3_1, F1_2 * F3_1 - F1_1 * F3_2, F1_1 * F2_2 - F1_2 * F2_1};
    R d2E[81];
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r) d2E[r + 9 * c] = g[r] * g[c];
    const R k = detF - a;
    auto sym = [&](int i, int j, R v) {
        d2E[i + 9 * j] += v;
        d2E[j + 9 * i] += v;
    };
    sym(0, 4, k * F3_3);  sym(0, 5, k * -F2_3); sym(0, 7, k * -F3_2); sym(0, 8, k * F2_2);
    sym(1, 3, k * -F3_3); sym(1, 5, k * F1_3);  sym(1, 6, k * F3_2);  sym(1, 8, k * -F1_2);
    sym(2, 3, k * F2_3);  sym(2, 4, k * -F1_3); sym(2, 6, k * -F2_2); sym(2, 7, k * F1_2);
    sym(3, 7, k * F3_1);  sym(3, 8, k * -F2_1);
    sym(4, 6, k * -F3_1); sym(4, 8, k * F1_1);
    sym(5, 6, k * F2_1);  sym(5, 7, k * -F1_1);
    for (int i = 0; i < 81; ++i) d2E[i] *= lmbd;
    for (int i = 0; i < 81; ++i) d2E[i] *= A;
    R dEdF[9];
    const R s = lmbd * (detF - a);
    for (int i = 0; i < 9; ++i) dEdF[i] = A * (miu * F[i] + s * g[i]);

    R m1, m2, m3;
    switch (order) {
    case 0:
        m1 = -DmInv[0] - DmInv[1] - DmInv[2];
        m2 = -DmInv[3] - DmInv[4] - DmInv[5];
        m3 = -DmInv[6] - DmInv[7] - DmInv[8];
        break;
    case 1: m1 = DmInv[0]; m2 = DmInv[3]; m3 = DmInv[6]; break;
    case 2: m1 = DmInv[1]; m2 = DmInv[4]; m3 = DmInv[7]; break;
    default: m1 = DmInv[2]; m2 = DmInv[5]; m3 = DmInv[8]; break;
    }
    R dEdx[3], h[9];
    assemble_vertex(dEdF, d2E, m1, m2, m3, dEdx, h);
    R dampingH[9];
    for (int i = 0; i < 9; ++i) dampingH[i] = h[i] * dampH;
    R tmp = (m1 * m1 + m2 * m2 + m3 * m3) * miu * A;
    h[0] += tmp; h[4] += tmp; h[8] += tmp;
    tmp *= dampD;
    dampingH[0] += tmp; dampingH[4] += tmp; dampingH[8] += tmp;
    for (int i = 0; i < 9; ++i) dampingH[i] /= dt;
    R df[3];
    for (int r = 0; r < 3; ++r) df[r] = dampingH[r] * disp[0] + dampingH[r + 3] * disp[1] + dampingH[r + 6] * disp[2];
    for (int r = 0; r < 3; ++r) force[r] -= dEdx[r] + df[r];
    for (int i = 0; i < 9; ++i) hess[i] += h[i] + dampingH[i];
}

// TetMeshFEM.cpp:275-325 (incidence; ascending tet id within each vertex).
template <class M> void build_incidence(M& m) {
    std::vector<int32_t> cnt(m.nv + 1, 0);
    for (int t = 0; t < m.nt; ++t)
        for (int k = 0; k < 4; ++k) cnt[m.tets[4 * t + k] + 1]++;
    m.incBegin.assign(m.nv + 1, 0);
    for (int i = 0; i < m.nv; ++i) m.incBegin[i + 1] = m.incBegin[i] + cnt[i + 1];
    m.incTet.resize(m.incBegin[m.nv]);
    m.incOrder.resize(m.incBegin[m.nv]);
    std::vector<int32_t> fill(m.incBegin.begin(), m.incBegin.end() - 1);
    for (int t = 0; t < m.nt; ++t)
        for (int k = 0; k < 4; ++k) {
            int v = m.tets[4 * t + k];
            m.incTet[fill[v]] = t;
            // first k with tetVIds(k) == v (TetMeshFEM.cpp:297-305)
            int ord = k;
            for (int kk = 0; kk < 4; ++kk)
                if (m.tets[4 * t + kk] == v) { ord = kk; break; }
            m.incOrder[fill[v]] = ord;
            fill[v]++;
        }
}


}  // namespace tetcore
