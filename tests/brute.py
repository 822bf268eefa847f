"""Dense weak-form assembly of the even-parity P1 operator, straight from
vertex coordinates (numpy), independent of the oracle and of the product:
  A[(i,o),(k,n)] = sum_t massfac_o delta_on kappa V/20 (1+delta_ik)
                 + sum_t (V/kappa) sum_ab d_a e_i d_b e_k E[a,b,o,n]
                 + sum_{vacuum F} S/12 (1+delta_ik) O[axis(F)][o,n]
  b[(i,0)]       = sum_t kappa V I_p / 4
diag=True keeps only a = b, o = n in the stiffness and o = n on the boundary
(the angular-diagonal operator of kernels.cu:290-357, 440-493)."""
import numpy as np


def boundary_faces(tets):
    faces = {}
    for t, tet in enumerate(tets):
        for j in range(4):
            key = tuple(sorted(int(tet[k]) for k in range(4) if k != j))
            faces.setdefault(key, []).append((t, j))
    return [(k, v[0]) for k, v in faces.items() if len(v) == 1]


def assemble(coords, tets, kappa, Ip, E, O, massfac, diag=False, face_ok=None):
    nP, slm = len(coords), E.shape[-1]
    N = nP * slm
    A = np.zeros((N, N))
    b = np.zeros(N)
    if diag:
        E = np.array([[np.diag(np.diag(E[a, a])) if a == c else np.zeros_like(E[a, c])
                       for c in range(3)] for a in range(3)])
        O = np.array([np.diag(np.diag(On)) for On in O])
    Mf = np.diag(massfac)
    for t, tet in enumerate(tets):
        P = coords[tet]
        J = (P[1:] - P[0]).T
        V = abs(np.linalg.det(J)) / 6
        g = np.linalg.inv(J)            # rows: grad e_1..3
        G = np.vstack([-g.sum(0), g])   # grad e_0..3
        for i in range(4):
            b[tet[i] * slm] += kappa[t] * V * Ip[t] / 4
            for k in range(4):
                blk = kappa[t] * V / 20 * (2 if i == k else 1) * Mf
                blk = blk + V / kappa[t] * np.einsum("a,b,abon->on", G[i], G[k], E)
                A[tet[i] * slm:(tet[i] + 1) * slm, tet[k] * slm:(tet[k] + 1) * slm] += blk
    for key, (t, j) in boundary_faces(tets):
        if face_ok is not None and not face_ok(coords[list(key)]):
            continue
        P = coords[list(key)]
        n = np.cross(P[1] - P[0], P[2] - P[0])
        S = np.linalg.norm(n) / 2
        axis = int(np.argmax(np.abs(n)))
        for i in range(3):
            for k in range(3):
                A[key[i] * slm:(key[i] + 1) * slm, key[k] * slm:(key[k] + 1) * slm] += \
                    S / 12 * (2 if i == k else 1) * O[axis]
    return A, b
