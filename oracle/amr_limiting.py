"""ORACLE -- test infrastructure only (tests, the smoke check and bench.py's
CPU-baseline leg may use it; the product never does).

A restatement of the reference's limiting ADER-DG step on a static adaptive
mesh, driven cell by cell through the C oracle's per-cell kernels
(oracle/*.c).  It follows, in the order of a non-fused step
(E/runners/Runner.cpp:1123-1140) with the limiter's local recomputation
(Runner.cpp:886-955, E/mappings/LocalRecomputation.cpp:170-210):

  Prediction   spaceTimePredictor + volumeIntegral on every leaf
               (E/solvers/ADERDGSolver.cpp:1818-1946)
  Sending      face data of a coarse leaf prolongated to the virtual fine
               cells on its boundary (prolongateFaceDataToDescendant,
               ADERDGSolver.cpp:2273-2345) and face data of the fine leaves
               restricted, after zeroing, into the refined coarse-level cell
               (prepareFaceDataOfAncestor, restrictToTopMostParent,
               ADERDGSolver.cpp:2240-2271, :2388-2440); DMP observables
               copied down and reduced up the same way (:2320-2345, :2442-2470)
  Merging      facewise limiter status (ADERDGSolver.cpp:2486-2545), min/max
               on every face (LimitingADERDGSolver.cpp:1816-1850), the Rusanov
               problem of every face that the limiter status allows and the
               FV ghost layers between limiter patches
               (mergeNeighboursBasedOnLimiterStatus,
               LimitingADERDGSolver.cpp:1641-1790; ghostLayerFilling,
               kernels/finitevolumes/commons/c/commons.cpph:876-982)
  Update       updateSolution's status switch (LimitingADERDGSolver.cpp:721-832),
               then DMP + admissibility and the new status
               (:838-1048); on an irregular limiter-domain change, status
               spreading (LimiterStatusSpreading, :369-379, five rounds as
               Runner.cpp:897-902), reinitialisation/rollback (:1283-1363),
               recomputation (:1378-1441) and min/max of every cell
               (:1050-1151)
  TimeStep     the ADER-DG stable time step of every leaf, global minimum.

What the mesh traversal cannot give here is restated on a static mesh:

* the mesh is built once (base n0^3 cells, periodic, refined by 3 where a
  cell box meets a spherical shell, 2:1 balanced) and never changes: a
  troubled cell below the finest level, or a limiter status above Ok in a
  finest cell next to a coarser cell, is counted as the mesh-update request
  the reference would make (LimitingADERDGSolver.cpp:953-955, :1516-1521)
  and the cell keeps its ADER-DG update;
* merges run in Jacobi order (every face of a sweep sees the values from
  before the sweep), which is what Peano's vertex-before-cell traversal
  gives for data that one face owns;
* an FV ghost layer whose neighbour holds no limiter patch (only possible
  next to a coarser cell, where the reference would refine) mirrors the
  patch's own boundary layers.

Limiter status values: Ok 0, NeighbourOfTroubled4 1, ..3 2, ..2 3, ..1 4,
Troubled 5 (E/ADERDGCellDescription.def:423-453).
"""
import itertools

import numpy as np

from . import oracle as O

OK, NOT4, NOT3, NOT2, NOT1, TROUBLED = range(6)
FV_STATUSES = (NOT2, NOT1, TROUBLED)


def dg_merge(s1, s2):
    """Rusanov problem between two finest-level cells of these statuses
    (mergeNeighboursBasedOnLimiterStatus, LimitingADERDGSolver.cpp:1695-1763)."""
    dg = (OK, NOT3, NOT4)
    if s1 in dg and s2 in dg:
        return True
    return (s1 in (NOT1, NOT2) and s2 in (NOT3, NOT4)) or (s2 in (NOT1, NOT2) and s1 in (NOT3, NOT4))


def fv_merge(s1, s2):
    """Exchange of FV ghost layers between the two cells' limiter patches (same place)."""
    if s1 in FV_STATUSES and s2 in FV_STATUSES:
        return True
    return (s1 in (NOT1, NOT2) and s2 in (NOT3, NOT4)) or (s2 in (NOT1, NOT2) and s1 in (NOT3, NOT4))


def can_communicate(s1, s2):
    """neighboursCanCommunicate (LimitingADERDGSolver.cpp:1682-1686); the
    reference aborts otherwise."""
    return (s1 > OK or s2 <= NOT3) and (s2 > OK or s1 <= NOT3)


GHOST = 2  # MUSCL-Hancock ghost width (Toolkit GenericFiniteVolumesMUSCLHancockInC.java:13)


# ---------------------------------------------------------------------------
# the static adaptive mesh

class Mesh:
    """Leaves of a 3-level (or fewer) spacetree over the unit cube.

    Level l has cell size 1 / (n0 3^l).  A present cell at level l < levels-1
    is refined when its box meets the shell r0 - w_l <= |x - c| <= r0 + w_l
    (w_l = widths[l]); refinement repeats with 2:1 balancing (no leaf has a
    face neighbour leaf two levels finer) until nothing changes.  Leaves are
    ordered by (level, z, y, x)."""

    def __init__(self, n0, levels, centre, radius, widths):
        self.n0, self.levels = n0, levels
        self.centre = np.asarray(centre, dtype=float)
        self.radius = float(radius)
        self.widths = list(widths) + [0.0] * levels
        self.refined = [set() for _ in range(levels)]
        changed = True
        while changed:
            changed = self._refine_by_shell() | self._balance()
        leaves = []
        for l in range(levels):
            for c in self._present(l):
                if c not in self.refined[l]:
                    leaves.append((l, c[2], c[1], c[0]))
        leaves.sort()
        self.leaves = [(l, (x, y, z)) for (l, z, y, x) in leaves]
        self.index = {lc: q for q, lc in enumerate(self.leaves)}
        self._faces()

    def n(self, l):
        return self.n0 * 3 ** l

    def h(self, l):
        return 1.0 / self.n(l)

    def _present(self, l):
        if l == 0:
            n = self.n0
            return [(i, j, k) for k in range(n) for j in range(n) for i in range(n)]
        out = []
        for p in sorted(self.refined[l - 1]):
            for a, b, c in itertools.product(range(3), repeat=3):
                out.append((3 * p[0] + a, 3 * p[1] + b, 3 * p[2] + c))
        return out

    def is_present(self, l, c):
        if l == 0:
            return True
        return (c[0] // 3, c[1] // 3, c[2] // 3) in self.refined[l - 1]

    def _meets_shell(self, l, c):
        h = self.h(l)
        lo = np.array(c, dtype=float) * h
        hi = lo + h
        near = np.clip(self.centre, lo, hi)
        dmin = np.linalg.norm(near - self.centre)
        far = np.where(np.abs(lo - self.centre) > np.abs(hi - self.centre), lo, hi)
        dmax = np.linalg.norm(far - self.centre)
        w = self.widths[l]
        return dmin <= self.radius + w and dmax >= self.radius - w

    def _refine_by_shell(self):
        changed = False
        for l in range(self.levels - 1):
            for c in self._present(l):
                if c not in self.refined[l] and self._meets_shell(l, c):
                    self.refined[l].add(c)
                    changed = True
        return changed

    def wrap(self, l, c):
        n = self.n(l)
        return (c[0] % n, c[1] % n, c[2] % n)

    def _balance(self):
        changed = False
        for l in range(self.levels - 2):
            for c in self._present(l):
                if c in self.refined[l]:
                    continue
                for d, s in itertools.product(range(3), (0, 1)):
                    nb = list(c)
                    nb[d] += 1 if s else -1
                    nb = self.wrap(l, nb)
                    if nb not in self.refined[l]:
                        continue
                    # children of nb on the shared face
                    hit = False
                    for a, b in itertools.product(range(3), repeat=2):
                        ch = [0, 0, 0]
                        t = [x for x in range(3) if x != d]
                        ch[d] = 3 * nb[d] + (0 if s else 2)
                        ch[t[0]] = 3 * nb[t[0]] + a
                        ch[t[1]] = 3 * nb[t[1]] + b
                        if tuple(ch) in self.refined[l + 1]:
                            hit = True
                    if hit:
                        self.refined[l].add(c)
                        changed = True
                        break
        return changed

    def _faces(self):
        """same: (left leaf, right leaf, d) -- left face 2d+1 meets right face 2d.
        ancestors: coarse leaf C, its face fc, the 9 fine leaves on the other
        side with their face index and subface index.  descendants: fine leaf
        F, its face ff, the coarse leaf C' and its face, subface index."""
        self.same, self.anc, self.desc = [], [], []
        self.nbr = np.full((len(self.leaves), 6), -1, dtype=np.int64)  # same-level leaf neighbours
        for q, (l, c) in enumerate(self.leaves):
            for d, s in itertools.product(range(3), (0, 1)):
                f = 2 * d + s
                nb = list(c)
                nb[d] += 1 if s else -1
                nb = self.wrap(l, nb)
                t = [x for x in range(3) if x != d]
                if self.is_present(l, nb):
                    if nb not in self.refined[l]:
                        r = self.index[(l, nb)]
                        self.nbr[q, f] = r
                        if s == 1:
                            self.same.append((q, r, d))
                    else:
                        kids = []
                        for a, b in itertools.product(range(3), repeat=2):
                            ch = [0, 0, 0]
                            ch[d] = 3 * nb[d] + (0 if s else 2)
                            ch[t[0]] = 3 * nb[t[0]] + a
                            ch[t[1]] = 3 * nb[t[1]] + b
                            sub = (ch[t[0]] - 3 * nb[t[0]], ch[t[1]] - 3 * nb[t[1]])
                            kids.append((self.index[(l + 1, tuple(ch))], 2 * d + (1 - s), sub))
                        self.anc.append((q, f, kids))
                else:
                    par = (nb[0] // 3, nb[1] // 3, nb[2] // 3)
                    r = self.index[(l - 1, par)]
                    sub = (nb[t[0]] - 3 * par[t[0]], nb[t[1]] - 3 * par[t[1]])
                    self.desc.append((q, f, r, 2 * d + (1 - s), sub))


# ---------------------------------------------------------------------------
# the limiting ADER-DG step

def _cell_layout(B):
    return B ** 3


class LimitingGrid:
    """Limiting ADER-DG (3D Euler, nobs = 5 DMP observables, MUSCL-Hancock
    subcells) on a static Mesh.  luh[q]: reference layout [z][y][x][5]."""

    def __init__(self, mesh, order, gamma=1.4, cfl=0.9, relax=1e-4, scaling=1e-3, steps_till_cured=0):
        self.m = mesh
        self.N = order
        self.B = order + 1
        self.P = 2 * order + 1
        self.S = self.P + 2 * GHOST
        self.c = O.cfg(order, gamma=gamma, cfl=cfl)
        self.relax, self.scaling, self.steps = relax, scaling, steps_till_cured
        nc = len(mesh.leaves)
        self.nc = nc
        self.finest = mesh.levels - 1
        self.level = np.array([l for (l, _) in mesh.leaves])
        self.luh = np.zeros((nc, self.B ** 3 * 5))
        self.status = np.zeros(nc, dtype=np.int64)
        self.fw = np.zeros((nc, 6), dtype=np.int64)
        self.cure = np.zeros(nc, dtype=np.int64)
        self.fmin = np.zeros((nc, 6, 5))
        self.fmax = np.zeros((nc, 6, 5))
        self.patch = {}  # cell -> [S^3 * 5] reference layout, ghost width 2
        self.t = 0.0
        self.counts = dict(irregular_steps=0, refinement_requests=0, coarse_fine_limiter=0, mirrored_ghost_faces=0,
                           cannot_communicate=0)

    def dx(self, q):
        h = self.m.h(self.level[q])
        return np.array([h, h, h])

    def centre(self, q):
        l, c = self.m.leaves[q]
        h = self.m.h(l)
        return (np.array(c, dtype=float) + 0.5) * h

    # -- initial conditions (setInitialConditions + the status after them,
    #    LimitingADERDGSolver.cpp:692-715, :906-930: admissibility only); a
    #    cell that holds a limiter patch gets the FV initial conditions at its
    #    subcell centres (kernels/finitevolumes/commons/c/commons.cpph:466-500)
    def init(self, ic):
        for q in range(self.nc):
            self.luh[q] = O.solution_adjustment(self.c, ic, self.centre(q), self.dx(q))
        self._start(ic)

    def init_from_luh(self, luh):
        """A given solution: patches are the projections of it."""
        self.luh[:] = np.asarray(luh, dtype=float).reshape(self.luh.shape)
        self._start(None)

    def _start(self, ic):
        self.status[:] = OK
        self.cure[:] = 0
        for q in range(self.nc):
            mn, mx = O.cell_min_max(self.N, self.luh[q])
            self.fmin[q, :] = mn
            self.fmax[q, :] = mx
            if self.level[q] == self.finest and not self._pad(mn):
                self.status[q] = TROUBLED
                self.cure[q] = self.steps + 1
            self.fw[q, :] = self.status[q]
        # patches of the limiter domain (ensureRequiredLimiterPatchIsAllocated)
        self._spread(5) if (self.status == TROUBLED).any() else None
        self.patch = {}
        for q in range(self.nc):
            if self.status[q] >= NOT4:
                if ic is None:
                    self.patch[q] = O.project_on_fv(self.N, self.luh[q], GHOST)
                else:
                    self.patch[q] = O.fv_patch_ic(self.c, ic, self.P, GHOST, self.centre(q), self.dx(q))

    def stable_dt(self):
        return min(O.stable_time_step_size(self.c, self.luh[q], self.dx(q)) for q in range(self.nc))

    @staticmethod
    def _pad(mn):
        """isPhysicallyAdmissible of the Euler solver: positive density and energy minima."""
        return mn[0] > 0.0 and mn[4] > 0.0

    # -- status machinery
    def _merge_facewise(self):
        fw = self.fw.copy()
        for (a, b, d) in self.m.same:
            if self.level[a] != self.finest:
                continue
            sa, sb = self.fw[a, 2 * d + 1], self.fw[b, 2 * d]
            fw[a, 2 * d + 1] = max(0, max(min(sa, TROUBLED), min(sb, TROUBLED)) - 1)
            fw[b, 2 * d] = max(0, max(min(sb, TROUBLED), min(sa, TROUBLED)) - 1)
        # faces of finest cells without a same-level partner merge with an Ok neighbour
        for q in np.nonzero(self.level == self.finest)[0]:
            for f in range(6):
                if self.m.nbr[q, f] < 0:
                    fw[q, f] = max(0, min(self.fw[q, f], TROUBLED) - 1)
        self.fw = fw

    def _spread(self, rounds):
        fin = self.level == self.finest
        for _ in range(rounds):
            self._merge_facewise()
            for q in np.nonzero(fin)[0]:
                if self.status[q] >= TROUBLED:
                    self.fw[q, :] = self.status[q]
                self.status[q] = self.fw[q].max()
                self.fw[q, :] = self.status[q]

    def _merge_minmax(self):
        """Face blocks after Sending + Merging: same-level faces take the
        min/max of both blocks; a fine face next to a coarser leaf merges
        with that leaf's face block (prolongated copy); a coarse face next to
        refined cells merges with the extremum of the nine fine face blocks."""
        mn, mx = self.fmin.copy(), self.fmax.copy()
        for (a, b, d) in self.m.same:
            lo = np.minimum(self.fmin[a, 2 * d + 1], self.fmin[b, 2 * d])
            hi = np.maximum(self.fmax[a, 2 * d + 1], self.fmax[b, 2 * d])
            mn[a, 2 * d + 1] = mn[b, 2 * d] = lo
            mx[a, 2 * d + 1] = mx[b, 2 * d] = hi
        for (q, f, r, fr, sub) in self.m.desc:
            mn[q, f] = np.minimum(self.fmin[q, f], self.fmin[r, fr])
            mx[q, f] = np.maximum(self.fmax[q, f], self.fmax[r, fr])
        for (q, f, kids) in self.m.anc:
            lo = np.full(5, np.finfo(float).max)
            hi = np.full(5, -np.finfo(float).max)
            for (k, fk, sub) in kids:
                lo = np.minimum(self.fmin[k, fk], lo)
                hi = np.maximum(self.fmax[k, fk], hi)
            mn[q, f] = np.minimum(self.fmin[q, f], lo)
            mx[q, f] = np.maximum(self.fmax[q, f], hi)
        self.fmin, self.fmax = mn, mx

    # -- FV patches
    def _limiter_min_max(self, q):
        S, G, P = self.S, GHOST, self.P
        a = self.patch_new[q].reshape(S, S, S, 5)[G:G + P, G:G + P, G:G + P]
        return a.reshape(-1, 5).min(axis=0), a.reshape(-1, 5).max(axis=0)

    def _fill_ghosts(self, q, patches, status):
        """ghostLayerFilling of cell q's patch from its six same-level
        neighbours' patch interiors (commons.cpph:876-982) where fv_merge
        allows the exchange."""
        S, G, P = self.S, GHOST, self.P
        me = patches[q].reshape(S, S, S, 5)
        lo, hi = G, G + P
        for d, s in itertools.product(range(3), (0, 1)):
            r = self.m.nbr[q, 2 * d + s]
            ax = 2 - d  # array axis of direction d ([z][y][x])
            for k in range(G):
                dst = [slice(lo, hi)] * 3
                src = [slice(lo, hi)] * 3
                dst[ax] = hi + k if s else k
                if r >= 0 and r in patches and fv_merge(status[q], status[r]):
                    nb = patches[r].reshape(S, S, S, 5)
                    src[ax] = lo + k if s else hi - G + k
                    me[tuple(dst)] = nb[tuple(src)]
                else:
                    # mirrored own boundary layers
                    src[ax] = hi - 1 - k if s else lo + G - 1 - k
                    me[tuple(dst)] = me[tuple(src)]
                    if k == 0:
                        self.counts["mirrored_ghost_faces"] += 1

    def _fv_update(self, q, dt):
        fresh = self.patch_old[q].copy()
        new, _ = O.musclhancock_update(self.c, self.P, fresh, fresh, self.dx(q), dt)
        self.patch_new[q] = new
        self.luh[q] = O.project_on_dg(self.N, new, GHOST)

    # -- one time step
    def step(self, dt):
        N, B, c = self.N, self.B, self.c
        B2 = B * B
        nc = self.nc
        fin = self.level == self.finest
        s_old = self.status.copy()
        # Prediction
        qf = np.zeros((nc, 6, B2 * 5))
        ff = np.zeros((nc, 6, B2 * 5))
        lduh = np.zeros((nc, B ** 3 * 5))
        for q in range(nc):
            p = O.space_time_predictor(c, self.luh[q], self.dx(q), dt)
            qf[q] = p["lQhbnd"].reshape(6, -1)
            ff[q] = p["lFhbnd"].reshape(6, -1)
            lduh[q] = O.volume_integral(c, p["lFhi"], self.dx(q))
        # Sending: face data of virtual cells (before any Riemann problem)
        desc_q, desc_f = [], []
        for (q, f, r, fr, sub) in self.m.desc:
            a, b = O.face_unknowns_prolongation(N, qf[r, fr], ff[r, fr], 0, 1, sub)
            desc_q.append(a)
            desc_f.append(b)
        anc_q, anc_f = [], []
        for (q, f, kids) in self.m.anc:
            sq, sf = np.zeros(B2 * 5), np.zeros(B2 * 5)
            for (k, fk, sub) in kids:
                sq, sf = O.face_unknowns_restriction(N, sq, sf, qf[k, fk], ff[k, fk], 0, 1, sub)
            anc_q.append(sq)
            anc_f.append(sf)
        # Merging: metadata, min/max, Riemann problems, limiter ghost layers
        self._merge_facewise()
        self._merge_minmax()
        for (a, b, d) in self.m.same:
            if self.level[a] == self.finest:
                if not can_communicate(s_old[a], s_old[b]):
                    self.counts["cannot_communicate"] += 1
                if not dg_merge(s_old[a], s_old[b]):
                    continue
            ff[a, 2 * d + 1], ff[b, 2 * d] = O.riemann_solver(c, ff[a, 2 * d + 1], ff[b, 2 * d], qf[a, 2 * d + 1],
                                                              qf[b, 2 * d], dt, d)
        for i, (q, f, r, fr, sub) in enumerate(self.m.desc):
            d = f // 2
            if f % 2:
                ff[q, f], _ = O.riemann_solver(c, ff[q, f], desc_f[i], qf[q, f], desc_q[i], dt, d)
            else:
                _, ff[q, f] = O.riemann_solver(c, desc_f[i], ff[q, f], desc_q[i], qf[q, f], dt, d)
        for i, (q, f, kids) in enumerate(self.m.anc):
            d = f // 2
            if f % 2:
                ff[q, f], _ = O.riemann_solver(c, ff[q, f], anc_f[i], qf[q, f], anc_q[i], dt, d)
            else:
                _, ff[q, f] = O.riemann_solver(c, anc_f[i], ff[q, f], anc_q[i], qf[q, f], dt, d)
        self.patch_old = {q: p.copy() for q, p in self.patch.items()}
        self.patch_new = {}
        for q in self.patch_old:
            if s_old[q] in FV_STATUSES:
                self._fill_ghosts(q, self.patch_old, s_old)
        # Update
        luh_prev = self.luh.copy()
        fv_done = np.zeros(nc, dtype=bool)
        for q in range(nc):
            if fin[q] and s_old[q] in FV_STATUSES:
                self._fv_update(q, dt)
                fv_done[q] = True
            else:
                lduh_q = O.surface_integral(c, lduh[q], ff[q].ravel(), self.dx(q))
                self.luh[q] = O.solution_update(c, self.luh[q], lduh_q, dt)
        # DMP + admissibility, new status (updateLimiterStatusAndMinAndMaxAfterSolutionUpdate)
        irregular = False
        for q in range(nc):
            mn, mx = O.cell_min_max(N, self.luh[q])
            bmin, bmax = self.fmin[q].min(axis=0), self.fmax[q].max(axis=0)
            sd = np.maximum((bmax - bmin) * self.scaling, self.relax)
            dmp_ok = not (np.any(mn < bmin - sd) or np.any(mx > bmax + sd))
            self.fmin[q, 0], self.fmax[q, 0] = mn, mx
            self.fmin[q, 1], self.fmax[q, 1] = bmin, bmax
            troubled = not (dmp_ok and self._pad(mn))
            if not fin[q]:
                if troubled:
                    self.counts["refinement_requests"] += 1
                continue
            if s_old[q] in FV_STATUSES:
                lmn, lmx = self._limiter_min_max(q)
                self.fmin[q, :] = lmn
                self.fmax[q, :] = lmx
            if troubled:
                self.cure[q] = self.steps + 1
                if s_old[q] not in (TROUBLED, NOT1):
                    irregular = True
                self.status[q] = TROUBLED
                self.fw[q, :] = TROUBLED
            elif s_old[q] == TROUBLED:
                self.status[q] = TROUBLED
                self.fw[q, :] = TROUBLED
                self.cure[q] -= 1
                if self.cure[q] == 0:
                    self.status[q] = NOT1
                    self.fw[q, :] = NOT1
                    self.cure[q] = self.steps
            self.status[q] = self.fw[q].max()
            self.fw[q, :] = self.status[q]
        if irregular:
            self.counts["irregular_steps"] += 1
            self._spread(5)
            for q in np.nonzero(fin & (self.status == TROUBLED))[0]:
                self.cure[q] = 1 + self.steps
            # rollback (reinitialiseSolvers): every patch returns to its value
            # at the start of the step; a cell without one (previously Ok)
            # projects its previous solution
            for q in np.nonzero(fin & (self.status >= NOT4))[0]:
                if q not in self.patch_old:
                    self.patch_old[q] = O.project_on_fv(N, luh_prev[q], GHOST)
            for q in np.nonzero(fin & np.isin(self.status, FV_STATUSES))[0]:
                self._fill_ghosts(q, self.patch_old, self.status)
            for q in np.nonzero(fin & np.isin(self.status, FV_STATUSES))[0]:
                if not fv_done[q]:
                    self.luh[q] = luh_prev[q]
                self._fv_update(q, dt)
                fv_done[q] = True
        # patches of the new limiter domain; a DG-updated cell's patch holds
        # its new solution
        self.patch = {}
        for q in np.nonzero(fin & (self.status >= NOT4))[0]:
            if fv_done[q]:
                self.patch[q] = self.patch_new[q]
            else:
                self.patch[q] = O.project_on_fv(N, self.luh[q], GHOST)
        if irregular:
            for q in range(nc):
                if fin[q] and self.status[q] in FV_STATUSES:
                    lmn, lmx = self._limiter_min_max(q)
                else:
                    lmn, lmx = O.cell_min_max(N, self.luh[q])
                self.fmin[q, :] = lmn
                self.fmax[q, :] = lmx
        for q in np.nonzero(fin)[0]:
            if self.status[q] > OK and (self.m.nbr[q] < 0).any():
                self.counts["coarse_fine_limiter"] += 1
        self.t += dt
        return self.stable_dt()
