"""CPU oracle of OWT-UCM (test infrastructure only: imported by tests/ and
nothing shipped).  A plain restatement of the reference's Contours2Ucm flow
(src/contours2ucm.cpp:405-490), step by step in the reference's own order, in
Python loops for small images:

  finest partition     CreateFinestPartition, src/contours2ucm.cpp:251-404
  watershed            watershed_yupeng, src/watershed.cpp:28-108
  contours             the reference's own lib_image contour_set, compiled from
                       its tree into oracle/_ref/libref_libimage.so
                       (ref_contours_driver.cc), as MexContourSides calls it
  doubled grid         SuperContour4C, src/contours2ucm.cpp:24-78
  cleaning             CleanWatersheds, :80-189, with bwmorph 'clean',
                       regionprops and bwlabel (src/MATLAB_functions.cpp:8-222)
  merging              ComputeUcm / UcmMeanPb, src/ucm_mean_pb.cpp:254-449
  calibration          NormalizeOutput, src/contours2ucm.cpp:191-249

Arrays are row-major numpy arrays (rows, cols) of the reference's cv::Mat
values.  Where the reference reads outside a matrix (CleanWatersheds at the
second row or column) the read is taken at the row-major flat position it
lands on, and as 0 before the first element.
"""
from __future__ import annotations

import ctypes as C
import heapq
import math
from collections import deque
from pathlib import Path

import numpy as np

_REF = Path(__file__).resolve().parent / "_ref" / "libref_libimage.so"
_lib = None


def _ref():
    global _lib
    if _lib is None:
        _lib = C.CDLL(str(_REF))
    return _lib


def ref_contours(bw: np.ndarray):
    """The reference's contours of a binary image: (vertices [nv, 2] (row,
    col), edges [ne, 2] (start, end), offsets [ne + 1], points [np, 2]); None
    where the reference throws."""
    bw = np.ascontiguousarray(bw != 0, np.uint8)
    H, W = bw.shape
    lib = _ref()
    P = C.c_void_p
    nv, npts = C.c_int(), C.c_longlong()
    ne = lib.ref_contours_run(bw.ctypes.data_as(P), W, H, C.byref(nv), C.byref(npts))
    if ne < 0:
        return None
    V = np.zeros(2 * max(nv.value, 1), np.int32)
    E = np.zeros(2 * max(ne, 1), np.int32)
    O = np.zeros(ne + 1, np.int64)
    Pt = np.zeros(2 * max(npts.value, 1), np.int32)
    lib.ref_contours_get(V.ctypes.data_as(P), E.ctypes.data_as(P), O.ctypes.data_as(P), Pt.ctypes.data_as(P))
    return (V[:2 * nv.value].reshape(-1, 2), E[:2 * ne].reshape(-1, 2), O, Pt[:2 * npts.value].reshape(-1, 2))


def watershed_yupeng(A: np.ndarray) -> np.ndarray:
    """src/watershed.cpp:28-108 with conn = 8."""
    rows, cols = A.shape
    L = np.full((rows, cols), -1, np.int64)
    couleur = 1
    for level in range(int(A.min()), int(A.max()) + 1):
        for x, y in np.argwhere(A == level):
            etat = -1
            for i in (-1, 0, 1):
                for j in (-1, 0, 1):
                    if i == 0 and j == 0:
                        continue
                    if 0 <= x + i < rows and 0 <= y + j < cols:
                        val = L[x + i, y + j]
                        if val > 0:
                            if etat == -1:
                                etat = val
                            elif val != etat:
                                etat = 0
            if etat >= 0:
                L[x, y] = etat
            else:
                L[x, y] = couleur
                couleur += 1
    return L


def pb_to_u8(pb: np.ndarray) -> np.ndarray:
    """pb.convertTo(CV_8UC1, alpha, beta) with the reference's alpha and beta
    (src/contours2ucm.cpp:270-278): round half to even, saturate."""
    lo, hi = pb.min(), pb.max()
    alpha = 255 / (hi - lo)
    beta = -255 * lo / (hi - lo)
    with np.errstate(invalid="ignore"):
        r = np.rint(pb * alpha + beta)
    out = np.where(np.isnan(r) | (r < 0), 0, np.minimum(r, 255))
    return out.astype(np.uint8)


def _orient(v1, v2) -> int:
    ang = np.float32(math.pi / 2)
    if v1[1] != v2[1]:
        ang = np.float32(math.atan((int(v1[0]) - int(v2[0])) * 1.0 / (int(v1[1]) - int(v2[1]))))
    deg = float(np.float32(float(np.float32(ang * np.float32(180))) / math.pi))
    if deg < -78.75 or deg >= 78.75:
        o = 1
    elif 56.25 <= deg < 78.75:
        o = 2
    elif 33.75 <= deg < 56.25:
        o = 3
    elif 11.25 <= deg < 33.75:
        o = 4
    elif -11.25 <= deg < 11.25:
        o = 5
    elif -33.75 <= deg < -11.25:
        o = 6
    elif -56.25 <= deg < -33.75:
        o = 7
    else:
        o = 8
    return o - 1


def create_finest_partition(pbo: np.ndarray):
    """src/contours2ucm.cpp:251-404: (ws_wt, pb8, ws, contours)."""
    pb = pbo[0].copy()
    for k in range(1, len(pbo)):
        pb = np.where(pb < pbo[k], pbo[k], pb)
    pb8 = pb_to_u8(pb)
    ws = watershed_yupeng(pb8)
    contours = ref_contours(ws == 0)
    ws_wt = np.zeros(pb.shape)
    if contours is None:
        return ws_wt, pb8, ws, contours
    V, E, O, Pt = contours
    for e in range(len(E)):
        v1, v2 = V[E[e, 0]], V[E[e, 1]]
        o = _orient(v1, v2)
        for k in range(O[e], O[e + 1]):
            r, c = Pt[k]
            ws_wt[r, c] = max(pbo[o][r, c], ws_wt[r, c])
        for r, c in (v1, v2):
            ws_wt[r, c] = max(pbo[o][r, c], ws_wt[r, c])
    return ws_wt, pb8, ws, contours


def super_contour_4c(pb: np.ndarray) -> np.ndarray:
    """src/contours2ucm.cpp:24-78."""
    tx, ty = pb.shape
    V = np.minimum(pb[:-1, :], pb[1:, :])
    Hm = np.minimum(pb[:, :-1], pb[:, 1:])
    rows, cols = 2 * tx, 2 * ty
    pb2 = np.zeros((rows, cols))
    for i in range(0, rows, 2):
        for j in range(0, cols, 2):
            pb2[i, j] = pb[i // 2, j // 2]
            if j != cols - 2:
                pb2[i, j + 1] = Hm[i // 2, j // 2]
            if i != rows - 2:
                pb2[i + 1, j] = V[i // 2, j // 2]
    pb2[rows - 1, :] = pb2[rows - 2, :]
    for i in range(rows):
        pb2[i, cols - 1] = max(pb2[i, cols - 1], pb2[i, cols - 2])
    return pb2


def bwmorph_clean(b: np.ndarray) -> np.ndarray:
    """bwmorph(.., 'clean', 8), src/MATLAB_functions.cpp:39-95: remove pixels
    with no 8-neighbour, repeated until nothing changes."""
    b = b.copy()
    rows, cols = b.shape
    while True:
        stable = True
        tmp = b.copy()
        for i in range(rows):
            for j in range(cols):
                if b[i, j] != 1:
                    continue
                iso = True
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        if (di or dj) and 0 <= i + di < rows and 0 <= j + dj < cols:
                            iso &= b[i + di, j + dj] == 0
                if iso:
                    tmp[i, j] = 0
                    stable = False
        b = tmp
        if stable:
            return b


def bwlabel(bw: np.ndarray) -> np.ndarray:
    """bwlabelManu with 8 neighbours, src/MATLAB_functions.cpp:182-217:
    breadth-first, components numbered from 1 in raster order."""
    m, n = bw.shape
    labels = np.zeros((m, n), np.int64)
    k = 0
    order = ((0, -1), (0, 1), (-1, -1), (-1, 0), (-1, 1), (1, -1), (1, 0), (1, 1))
    for i in range(m):
        for j in range(n):
            if bw[i, j] == 1 and labels[i, j] == 0:
                k += 1
                labels[i, j] = k
                q = deque([(i, j)])
                while q:
                    I, J = q.popleft()
                    for di, dj in order:
                        a, b = I + di, J + dj
                        if 0 <= a < m and 0 <= b < n and bw[a, b] == 1 and labels[a, b] == 0:
                            labels[a, b] = k
                            q.append((a, b))
    return labels


def clean_watersheds(ws: np.ndarray) -> np.ndarray:
    """src/contours2ucm.cpp:80-189 on a copy."""
    ws = ws.copy()
    rows, cols = ws.shape
    flat = ws.reshape(-1)
    c = bwmorph_clean((ws == 0).astype(np.uint8))
    artifacts = ((c == 0) & (ws == 0)).astype(np.uint8)
    lab = bwlabel(artifacts)
    firsts = {}
    for i in range(rows):
        for j in range(cols):
            if lab[i, j] > 0 and lab[i, j] not in firsts:
                firsts[lab[i, j]] = (i, j)

    def at(i, j):
        f = i * cols + j
        return 0.0 if f < 0 else flat[f]

    for r in range(1, lab.max() + 1):
        xc, yc = firsts[r]
        vec = [max(at(xc - 2, yc - 1), at(xc - 1, yc - 2)), max(at(xc + 2, yc - 1), at(xc + 1, yc - 2)),
               max(at(xc + 2, yc + 1), at(xc + 1, yc + 2)), max(at(xc - 2, yc + 1), at(xc - 1, yc + 2))]
        nd, idx = vec[0], 0
        for i in range(1, 4):
            if vec[i] < nd:
                idx, nd = i, vec[i]
        if idx == 0:
            if at(xc - 2, yc - 1) < at(xc - 1, yc - 2):
                ws[xc, yc - 1], ws[xc - 1, yc] = 0, vec[0]
            else:
                ws[xc, yc - 1], ws[xc - 1, yc] = vec[0], 0
            ws[xc - 1, yc - 1] = vec[0]
        elif idx == 1:
            if at(xc + 2, yc - 1) < at(xc + 1, yc - 2):
                ws[xc, yc - 1], ws[xc + 1, yc] = 0, vec[1]
            else:
                ws[xc, yc - 1], ws[xc + 1, yc] = vec[1], 0
            ws[xc + 1, yc - 1] = vec[1]
        elif idx == 2:
            if at(xc + 2, yc + 1) < at(xc + 1, yc + 2):
                ws[xc, yc + 1], ws[xc + 1, yc] = 0, vec[2]
            else:
                ws[xc, yc + 1], ws[xc + 1, yc] = vec[2], 0
            ws[xc + 1, yc + 1] = vec[2]
        else:
            if at(xc - 2, yc + 1) < at(xc - 1, yc + 2):
                ws[xc, yc + 1], ws[xc - 1, yc] = 0, vec[3]
            else:
                ws[xc, yc + 1], ws[xc - 1, yc] = vec[3], 0
            ws[xc - 1, yc + 1] = vec[3]
    return ws


def ucm_mean_pb(ws_wt2: np.ndarray, labels: np.ndarray) -> np.ndarray:
    """UcmMeanPb / ComputeUcm, src/ucm_mean_pb.cpp:254-449, in the reference's
    column-major indexing; returns the (2 fil + 1) x (2 col + 1) map."""
    fil, col = labels.shape
    txc, tyc = 2 * fil + 1, 2 * col + 1
    local = ws_wt2.reshape(-1, order="F")
    part = labels.reshape(-1, order="F")
    totcc = int(part.max()) + 1
    tx, ty = fil, col
    lab = list(range(totcc))
    ucm = np.zeros(txc * tyc)
    boundary = [[] for _ in range(totcc)]
    elements = [[c] for c in range(totcc)]
    neighbors = [dict() for _ in range(totcc)]  # n -> [energy, totalPb, bdryLength]
    vx, vy = (1, 0, -1, 0), (0, 1, 0, -1)
    for p in range(tx * ty):
        xp, yp = p % tx, p // tx
        for v in range(4):
            nxp, nyp = xp + vx[v], yp + vy[v]
            cnp = nxp + nyp * tx
            if 0 <= nyp < ty and 0 <= nxp < tx and part[cnp] != part[p]:
                boundary[part[p]].append(((xp + nxp + 1) + (yp + nyp + 1) * (2 * tx + 1), int(part[cnp])))
    queue = []
    for c in range(totcc):
        b = sorted(boundary[c], key=lambda e: (e[1], e[0]))
        boundary[c] = b
        if not b:
            continue
        label = b[0][1]
        tot_b, tot_pb = 0.0, 0.0
        for coord, nb in b:
            if nb == label:
                tot_b += 1
                tot_pb += local[coord]
            else:
                neighbors[c][label] = [tot_pb / tot_b, tot_pb, tot_b]
                if label > c:
                    heapq.heappush(queue, (tot_pb / tot_b, c, label))
                label, tot_b, tot_pb = nb, 1.0, local[coord]
        neighbors[c][label] = [tot_pb / tot_b, tot_pb, tot_b]
        if label > c:
            heapq.heappush(queue, (tot_pb / tot_b, c, label))
    while queue:
        energy, r1, r2 = heapq.heappop(queue)
        if lab[r1] != r1 or lab[r2] != r2 or energy != neighbors[r1][r2][0]:
            continue
        sal = neighbors[r1][r2][1] / neighbors[r1][r2][2]
        son, father = (r1, r2) if r1 < r2 else (r2, r1)
        boundary[father] = [e for e in boundary[father] if lab[e[1]] != son]
        for coord, nb in boundary[son]:
            if ucm[coord] < sal:
                ucm[coord] = sal
            if lab[nb] != father:
                boundary[father].append((coord, nb))
        boundary[son] = []
        for e in elements[son]:
            lab[e] = father
        elements[father] = elements[son] + elements[father]
        elements[son] = []
        neighbors[father] = {n: v for n, v in neighbors[father].items() if lab[n] == n}
        neighbors[son] = {n: v for n, v in neighbors[son].items() if lab[n] == n and lab[n] != father}
        for n in sorted(neighbors[son]):
            v = neighbors[son][n]
            a = neighbors[father].setdefault(n, [0.0, 0.0, 0.0])
            b = neighbors[n].setdefault(father, [0.0, 0.0, 0.0])
            a[1] += v[1]
            b[1] += v[1]
            a[2] += v[2]
            b[2] += v[2]
        neighbors[son] = {}
        for n in sorted(neighbors[father]):
            d = neighbors[father][n][1] / neighbors[father][n][2]
            heapq.heappush(queue, (d, n, father))
            neighbors[father][n][0] = d
            neighbors[n].setdefault(father, [0.0, 0.0, 0.0])[0] = d
    for x in range(0, txc, 2):
        for y in range(0, tyc, 2):
            m = 0.0
            for v in range(4):
                nxp, nyp = x + vx[v], y + vy[v]
                if 0 <= nyp < tyc and 0 <= nxp < txc and m < ucm[nxp + nyp * txc]:
                    m = ucm[nxp + nyp * txc]
            ucm[x + y * txc] = m
    return ucm.reshape(txc, tyc, order="F")


def normalize_output(pb: np.ndarray) -> np.ndarray:
    """src/contours2ucm.cpp:191-249, elementwise with the C library's exp
    (math.exp), as the reference's std::exp (numpy's vectorised exp may round
    differently)."""
    out = np.empty(pb.shape)
    flat_in, flat_out = pb.reshape(-1), out.reshape(-1)
    for i in range(flat_in.size):
        comp = 1 * -2.7487 + float(flat_in[i]) * 11.1189
        p = 1.0 / (1.0 + math.exp(-comp))
        p = (p - 0.0602) / (1.0 - 0.0602)
        flat_out[i] = min(1.0, max(0.0, p))
    return out


def contours2ucm(pbo: np.ndarray, double_size: bool = True):
    """src/contours2ucm.cpp:405-490; returns (ucm, stages dict)."""
    ws_wt, pb8, ws, contours = create_finest_partition(pbo)
    pb2 = super_contour_4c(ws_wt)
    pb2 = clean_watersheds(pb2)
    rows, cols = pb2.shape
    labels2 = bwlabel((pb2 == 0).astype(np.uint8))
    labels = labels2[1::2, 1::2][: rows // 2, : cols // 2] - 1
    ws_wt2 = np.empty((rows + 1, cols + 1))
    ws_wt2[:rows, :cols] = pb2
    ws_wt2[rows, :cols] = pb2[rows - 1]
    ws_wt2[:, cols] = ws_wt2[:, cols - 1]
    super_ucm = ucm_mean_pb(ws_wt2, labels)
    norm = normalize_output(super_ucm)
    if double_size:
        ucm = norm
    else:
        R, Cc = norm.shape
        ucm = np.zeros(((R + 1) // 2, (Cc + 1) // 2))
        for i in range(2, R, 2):
            for j in range(2, Cc, 2):
                ucm[(i - 2) // 2, (j - 2) // 2] = norm[i, j]
    stages = dict(pb8=pb8, ws=ws, ws_wt=ws_wt, ws_wt2=ws_wt2, labels=labels, super_ucm=super_ucm,
                  regions=int(labels.max()) + 1, contours=contours)
    return ucm, stages
