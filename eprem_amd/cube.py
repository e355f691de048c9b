"""Cube-sphere stream topology: the four neighbours (north, east, west,
south on the cube surface) of every stream, for shell diffusion
(DiffuseShellData, src/energeticParticles.c:418-555).

Stream s = (face * rows + row) * cols + col sits at the centre of its cell
on the cube face (initCubeCoords, src/cubeShellInit.c:529-822).  A
neighbour is the cell one cell width away along the face: within the face
it is the adjacent (row, col); across a cube edge the displaced point is
folded over the edge onto the next face and the neighbour is the cell whose
centre is nearest to it (initNEWS, src/cubeShellInit.c:95-520, links the
same cells).  Only the set of links enters the diffusion result: each
link's weight depends on the two nodes it joins, and every target sums its
incoming terms in source order.
"""
import numpy as np

# (col axis, row axis) of each face in cube coordinates, per cell of width
# 1/cols (1/rows): TOP, BOTTOM, FRONT, BACK, LEFT, RIGHT
_COL_AXIS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0]], float)
_ROW_AXIS = np.array([[0, -1, 0], [0, -1, 0], [0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1]], float)


def cell_centres(rows, cols):
    """[S, 3] cell centres on the cube surface (|coordinate| = 0.5 on the
    face axis) and the face of each stream."""
    S = 6 * rows * cols
    P = np.zeros((S, 3))
    face_of = np.zeros(S, int)
    cw, ch = 1.0 / cols, 1.0 / rows
    for face in range(6):
        for row in range(rows):
            for col in range(cols):
                u = -0.5 + 0.5 * cw + col * cw
                u2 = 0.5 - 0.5 * cw - col * cw
                w = 0.5 - 0.5 * ch - row * ch
                P[(face * rows + row) * cols + col] = [(u, w, 0.5), (u2, w, -0.5), (0.5, u, w), (-0.5, u2, w),
                                                       (u2, 0.5, w), (u, -0.5, w)][face]
                face_of[(face * rows + row) * cols + col] = face
    return P, face_of


def _fold(p):
    """Move a point displaced past a cube edge onto the neighbouring face."""
    q = p.copy()
    a = np.abs(q)
    t = int(np.argmax(a))
    if a[t] <= 0.5 + 1e-12:
        return q
    over = a[t] - 0.5
    # the axis that was the face normal: the other coordinate at |0.5|
    n = [i for i in range(3) if i != t and abs(abs(q[i]) - 0.5) < 1e-9][0]
    q[t] = np.sign(q[t]) * 0.5
    q[n] = np.sign(q[n]) * (0.5 - over)
    return q


def neighbours(rows, cols):
    """[S, 4] stream indices of the north, east, west and south neighbours."""
    P, face_of = cell_centres(rows, cols)
    S = len(P)
    cw, ch = 1.0 / cols, 1.0 / rows
    out = np.zeros((S, 4), np.int32)
    for s in range(S):
        f = face_of[s]
        moves = (-ch * _ROW_AXIS[f], cw * _COL_AXIS[f], -cw * _COL_AXIS[f], ch * _ROW_AXIS[f])
        for d, mv in enumerate(moves):
            q = _fold(P[s] + mv)
            out[s, d] = int(np.argmin(np.sum((P - q) ** 2, axis=1)))
    return out
