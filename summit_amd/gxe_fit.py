"""Variance components G + GxE + NxE from the engine's GxE LD scores: the
method-of-moments normal equations and their solve, in FP64 on the host.

Restated from SUMMIT's summary-statistics GxE estimator
(src/summit/inference/gxe.py).  With r the rank of the residual maker P
(N - p - 1 for p covariates and the intercept), f_xj = P g_j and
f_wj = P(e . g_j) the SNP features normalized to f'f = r, and the phenotype
projected and scaled to y'y = r, the kernels are

    K_G,a = sum_j a_ja f_xj f_xj' / A_a,   K_GxE,a = sum_j a_ja f_wj f_wj' / A_a,
    K_NxE = P diag(e^2) P,                 K_res = P,

(A_a the mass of annotation a) and the normal equations are
sum_d tr(K_c K_d) s_d = y' K_c y.  From summary arrays (gxe.py:650-816):

* y' K_G,a y = r sum_j a_ja score_xj^2 / A_a with score = f'y / sqrt(r) (the
  GWAS and GWIS marginal scores), tr K_G,a = r sum_j a_ja norm_xj / A_a with
  norm = f'f / r, tr(K_G,a K_NxE) = r sum_j a_ja diag_nxe_xj / A_a with
  diag_nxe = f' diag(e^2) f / r;
* tr(K_G,a K_G,b) = r^2 sum_j a_ja ld_xx[j, b] / (A_a A_b) with the LD panels
  ld[j, b] = sum_i a_ib (f_j' f_i / r)^2 -- the engine's XX, XW, WX, WW
  directional scores -- averaged over both orientations;
* the NxE and residual rows: y' K_NxE y = |e . y|^2, y'y = r,
  tr K_NxE and tr K_NxE^2 exact (src/summit/ldscore/gxe_score.py:1213-1236).

The solve (gxe.py:819-997) eliminates the residual row
(s_res = 1 - t's / r with t the traces) and solves the reduced system with an
SVD pseudo-inverse, after checking symmetry, the diagonal, the Cauchy-Schwarz
bounds of a Gram matrix and positive semi-definiteness.  Delete-one block
jackknife (gxe.py:1159-1290, :1496-1522, :2036-2066): the LD panels stay
frozen, each deleted system drops the block's SNPs from the left-hand sums and
the annotation masses, and SE^2 = (J - 1)/J sum (theta_(j) - mean)^2.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import Sequence

import numpy as np


def _finite(name: str, value, ndim: int | None = None) -> np.ndarray:
    arr = np.asarray(value, dtype=np.float64)
    if ndim is not None and arr.ndim != ndim:
        raise ValueError(f"{name}: expected {ndim} dimensions, got shape {arr.shape}.")
    if not np.all(np.isfinite(arr)):
        raise ValueError(f"{name} has non-finite entries.")
    return arr


@dataclass(frozen=True)
class GxENormalEquations:
    matrix: np.ndarray        # (P, P) tr(K_c K_d)
    rhs: np.ndarray           # (P,) y' K_c y
    traces: np.ndarray        # (P,) tr K_c
    component_names: tuple
    diagnostics: dict | None = None


@dataclass(frozen=True)
class GxEFitResult:
    component_names: tuple
    coefficients: np.ndarray      # per-kernel variance parameters s_c
    contributions: np.ndarray     # s_c tr K_c
    proportions: np.ndarray       # contributions / their sum
    normal_matrix: np.ndarray
    rhs: np.ndarray
    singular_values: np.ndarray   # of the residual-eliminated matrix
    normal_eigenvalues: np.ndarray
    rank: int
    condition_number: float
    relative_residual: float
    solve_method: str = "residual_eliminated_svd"
    normal_symmetry_error: float = 0.0
    cauchy_schwarz_max_violation: float = 0.0
    component_influence: np.ndarray | None = None
    identifiable: bool = True
    coefficient_standard_errors: np.ndarray | None = None
    proportion_standard_errors: np.ndarray | None = None
    jackknife_coefficients: np.ndarray | None = None
    jackknife_proportions: np.ndarray | None = None


# --------------------------------------------------------------------------- assembly
def _check_inputs(annotations, vectors: dict, panels: dict):
    annot = _finite("annotations", annotations, ndim=2)
    m, k = annot.shape
    if m == 0 or k == 0:
        raise ValueError("annotations need at least one SNP and one column.")
    if np.any(annot < 0.0):
        raise ValueError("annotations must be non-negative.")
    vec = {}
    for name, value in vectors.items():
        a = _finite(name, value, ndim=1)
        if a.shape != (m,):
            raise ValueError(f"{name} has shape {a.shape}, expected ({m},).")
        if name in ("norm_x", "norm_w", "diag_nxe_x", "diag_nxe_w") and np.any(a < 0.0):
            raise ValueError(f"{name} must be non-negative (first negative row {int(np.flatnonzero(a < 0)[0])}).")
        vec[name] = a
    pan = {}
    for name, value in panels.items():
        a = _finite(name, value, ndim=2)
        if a.shape != (m, k):
            raise ValueError(f"{name} has shape {a.shape}, expected ({m}, {k}).")
        if np.any(a < 0.0):
            raise ValueError(f"{name} must hold non-negative squared-sketch sums "
                             f"(first negative entry {np.argwhere(a < 0.0)[0].tolist()}).")
        pan[name] = a
    masses = annot.sum(axis=0)
    if np.any(~np.isfinite(masses)) or np.any(masses <= 0.0):
        raise ValueError(f"annotation columns without positive mass: "
                         f"{np.flatnonzero(~np.isfinite(masses) | (masses <= 0.0)).tolist()}.")
    return annot, vec, pan


def _check_scalars(r, q_nxe, q_residual, trace_nxe, trace_nxe_sq):
    r = int(r)
    if r <= 0:
        raise ValueError(f"residual_rank must be positive, got {r}.")
    vals = [float(q_nxe), float(q_residual), float(trace_nxe), float(trace_nxe_sq)]
    if not all(np.isfinite(v) for v in vals):
        raise ValueError("the NxE and residual moments must be finite.")
    if vals[0] < 0.0 or vals[1] <= 0.0:
        raise ValueError("the phenotype moments need q_nxe >= 0 and q_residual > 0.")
    if not np.isclose(vals[1], float(r), rtol=1e-10, atol=1e-8):
        raise ValueError(f"q_residual = {vals[1]:.16g} does not match the normalized score scale "
                         f"(residual_rank = {r}).")
    if vals[2] < 0.0 or vals[3] < 0.0:
        raise ValueError("the NxE traces must be non-negative.")
    return r, vals


def _component_names(k: int, annotation_names):
    if annotation_names is None:
        names = tuple(f"bin{i}" for i in range(k))
    else:
        names = tuple(str(x) for x in annotation_names)
        if len(names) != k or len(set(names)) != k:
            raise ValueError("annotation_names needs one distinct name per annotation column.")
    return tuple(f"G:{n}" for n in names) + tuple(f"GxE:{n}" for n in names) + ("NxE", "residual")


def _system(k: int, r: float, mass, sx2, sw2, nx, nw, dx, dw, cross, q_nxe, q_res, t_nxe, t_nxe2):
    """Normal matrix, rhs and traces from annotation-weighted sums over the (kept)
    SNPs: sx2[a] = sum a_ja score_xj^2, nx/nw norms, dx/dw NxE diagonals, and
    cross[key][left, source] = sum_j a_j,left ld_key[j, source]."""
    p = 2 * k + 2
    lhs = np.zeros((p, p))
    rhs = np.zeros(p)
    tr = np.zeros(p)
    g, w = np.arange(k), k + np.arange(k)
    rhs[g] = r * sx2 / mass
    rhs[w] = r * sw2 / mass
    tr[g] = r * nx / mass
    tr[w] = r * nw / mass
    lhs[g, 2 * k] = lhs[2 * k, g] = r * dx / mass
    lhs[w, 2 * k] = lhs[2 * k, w] = r * dw / mass
    lhs[g, 2 * k + 1] = lhs[2 * k + 1, g] = tr[g]
    lhs[w, 2 * k + 1] = lhs[2 * k + 1, w] = tr[w]
    den = mass[:, None] * mass[None, :]
    xx = r * r * cross["ld_xx"] / den
    ww = r * r * cross["ld_ww"] / den
    # both orientations of each kernel pair, averaged
    lhs[:k, :k] = 0.5 * (xx + xx.T)
    lhs[k:2 * k, k:2 * k] = 0.5 * (ww + ww.T)
    gw = 0.5 * (r * r * cross["ld_xw"] / den + (r * r * cross["ld_wx"] / den).T)
    lhs[:k, k:2 * k] = gw
    lhs[k:2 * k, :k] = gw.T
    rhs[2 * k], rhs[2 * k + 1] = q_nxe, q_res
    tr[2 * k], tr[2 * k + 1] = t_nxe, r
    lhs[2 * k, 2 * k] = t_nxe2
    lhs[2 * k, 2 * k + 1] = lhs[2 * k + 1, 2 * k] = t_nxe
    lhs[2 * k + 1, 2 * k + 1] = r
    return 0.5 * (lhs + lhs.T), rhs, tr


def _cross_sums(annot, panels):
    return {key: annot.T @ p for key, p in panels.items()}


def assemble_normal_equations(*, annotations, score_x, score_w, ld_xx, ld_xw, ld_wx, ld_ww, norm_x, norm_w,
                              diag_nxe_x, diag_nxe_w, residual_rank, q_nxe, q_residual, trace_nxe, trace_nxe_sq,
                              annotation_names: Sequence[str] | None = None,
                              ld_scale: str = "cross_product_over_rank_squared") -> GxENormalEquations:
    """The 2K + 2 normal equations of G (K annotations), GxE (K), NxE and the
    residual from aligned per-SNP summary arrays (src/summit/inference/gxe.py:650-816).
    score_* = f'y / sqrt(r); ld_* = sum over source SNPs of a (f'f_i / r)^2 per
    source annotation; norm_* = f'f / r; diag_nxe_* = f' diag(e^2) f / r."""
    annot, vec, pan = _check_inputs(
        annotations, {"score_x": score_x, "score_w": score_w, "norm_x": norm_x, "norm_w": norm_w,
                      "diag_nxe_x": diag_nxe_x, "diag_nxe_w": diag_nxe_w},
        {"ld_xx": ld_xx, "ld_xw": ld_xw, "ld_wx": ld_wx, "ld_ww": ld_ww})
    r, (qn, qr, tn, tn2) = _check_scalars(residual_rank, q_nxe, q_residual, trace_nxe, trace_nxe_sq)
    if ld_scale != "cross_product_over_rank_squared":
        raise ValueError(f"unsupported LD scale {ld_scale!r}: the panels must hold (f'f_i / r)^2 sums.")
    k = annot.shape[1]
    names = _component_names(k, annotation_names)
    lhs, rhs, tr = _system(k, float(r), annot.sum(axis=0), annot.T @ vec["score_x"] ** 2,
                           annot.T @ vec["score_w"] ** 2, annot.T @ vec["norm_x"], annot.T @ vec["norm_w"],
                           annot.T @ vec["diag_nxe_x"], annot.T @ vec["diag_nxe_w"], _cross_sums(annot, pan),
                           qn, qr, tn, tn2)
    return GxENormalEquations(lhs, rhs, tr, names)


# --------------------------------------------------------------------------- solve
def solve_normal_equations(equations: GxENormalEquations, *, rcond: float | None = None,
                           max_condition: float = 1.0e12, allow_ill_conditioned: bool = False) -> GxEFitResult:
    """Residual-eliminated SVD solve of the normal equations (src/summit/inference/gxe.py:819-997)."""
    max_condition = float(max_condition)
    if not (np.isfinite(max_condition) and max_condition > 0.0):
        raise ValueError(f"max_condition must be positive and finite, got {max_condition!r}.")
    if rcond is not None:
        rcond = float(rcond)
        if not (np.isfinite(rcond) and rcond > 0.0):
            raise ValueError(f"rcond must be positive and finite, got {rcond!r}.")
    raw = _finite("normal matrix", equations.matrix, ndim=2)
    rhs = _finite("normal rhs", equations.rhs, ndim=1)
    tr = _finite("traces", equations.traces, ndim=1)
    n = rhs.shape[0]
    if raw.shape != (n, n) or n < 2 or tr.shape != (n,):
        raise ValueError("the normal matrix must be square and match the rhs and traces.")
    scale = max(float(np.max(np.abs(raw))), 1.0)
    sym_err = float(np.max(np.abs(raw - raw.T)))
    if sym_err > 2.0e-10 * scale:
        raise ValueError(f"the normal matrix is not symmetric: max |A - A'| = {sym_err:.6g}.")
    A = 0.5 * (raw + raw.T)
    diag = np.diag(A)
    if float(np.min(diag)) < -2.0e-10 * scale:
        raise ValueError(f"a kernel norm on the diagonal is negative: {float(np.min(diag)):.6g}.")
    bound = np.sqrt(np.maximum(diag, 0.0)[:, None] * np.maximum(diag, 0.0)[None, :])
    excess = np.abs(A) - bound
    cs_max = float(np.max(excess))
    if cs_max > 2.0e-10 * max(scale, float(np.max(bound)), 1.0):
        at = np.unravel_index(int(np.argmax(excess)), excess.shape)
        raise ValueError(f"the kernel Gram matrix breaks a Cauchy-Schwarz bound at {at} by {cs_max:.6g}: "
                         "it is not positive semi-definite.")
    eig = np.linalg.eigvalsh(A)
    if eig[0] < -1.0e-10 * max(float(np.max(np.abs(eig))), 1.0):
        raise ValueError(f"the normal matrix is not positive semi-definite (smallest eigenvalue {eig[0]:.6g}): "
                         "the LD panels are too noisy for this system; use more probes.")
    t_res = float(tr[-1])
    if t_res <= 0.0:
        raise ValueError("the residual kernel's trace must be positive.")
    tol_res = 2.0e-10 * max(scale, abs(t_res), 1.0)
    if (abs(A[-1, -1] - t_res) > tol_res or float(np.max(np.abs(A[:-1, -1] - tr[:-1]))) > tol_res
            or abs(rhs[-1] - t_res) > tol_res):
        raise ValueError("the residual row does not match the kernel traces and the normalized phenotype.")
    t = tr[:-1]
    red = A[:-1, :-1] - np.outer(t, t) / t_res
    red = 0.5 * (red + red.T)
    red_rhs = rhs[:-1] - t
    u, sv, vh = np.linalg.svd(red, full_matrices=False)
    if sv.size == 0 or sv[0] <= 0.0:
        raise ValueError("the residual-eliminated normal matrix has no positive singular value.")
    tol = max(red.shape) * np.finfo(np.float64).eps * sv[0] if rcond is None else rcond * sv[0]
    red_rank = int(np.sum(sv > tol))
    cond = float(np.inf if sv[-1] <= 0.0 else sv[0] / sv[-1])
    if not allow_ill_conditioned and (red_rank < red.shape[0] or cond > max_condition):
        raise ValueError(f"the residual-eliminated equations are not identifiable: rank {red_rank + 1}/{n}, "
                         f"condition {cond:.6g} (a constant e^2 makes NxE the residual; redundant annotations "
                         "make kernels collinear).")
    inv_sv = np.zeros_like(sv)
    inv_sv[sv > tol] = 1.0 / sv[sv > tol]
    pinv = (vh.T * inv_sv[None, :]) @ u.T
    coef_nr = pinv @ red_rhs
    coef = np.concatenate([coef_nr, [1.0 - float(t @ coef_nr) / t_res]])
    rel_res = float(np.linalg.norm(A @ coef - rhs) / max(float(np.linalg.norm(rhs)), np.finfo(np.float64).tiny))
    influence = np.empty(n)
    influence[:-1] = np.linalg.norm(pinv, axis=1)
    influence[-1] = math.sqrt(1.0 + float(np.linalg.norm(t @ pinv / t_res) ** 2))
    contrib = coef * tr
    total = float(contrib.sum())
    prop = np.full_like(contrib, np.nan)
    if np.isfinite(total) and abs(total) > np.finfo(np.float64).tiny:
        prop = contrib / total
    return GxEFitResult(component_names=tuple(equations.component_names), coefficients=coef, contributions=contrib,
                        proportions=prop, normal_matrix=A, rhs=rhs, singular_values=sv, normal_eigenvalues=eig,
                        rank=red_rank + 1, condition_number=cond, relative_residual=rel_res,
                        normal_symmetry_error=sym_err, cauchy_schwarz_max_violation=max(0.0, cs_max),
                        component_influence=influence,
                        identifiable=bool(red_rank == red.shape[0] and cond <= max_condition))


# --------------------------------------------------------------------------- jackknife
def deleted_block_equations(equations_inputs: dict, blocks) -> list:
    """One normal system per deleted block (contiguous integer block ids 0..J-1
    per SNP): the frozen LD panels with the block's SNPs dropped from every
    left-hand annotation sum and from the masses on both sides
    (src/summit/inference/gxe.py:1159-1290, :1496-1522)."""
    d = equations_inputs
    annot, vec, pan = _check_inputs(
        d["annotations"], {key: d[key] for key in ("score_x", "score_w", "norm_x", "norm_w", "diag_nxe_x",
                                                    "diag_nxe_w")},
        {key: d[key] for key in ("ld_xx", "ld_xw", "ld_wx", "ld_ww")})
    r, (qn, qr, tn, tn2) = _check_scalars(d["residual_rank"], d["q_nxe"], d["q_residual"], d["trace_nxe"],
                                          d["trace_nxe_sq"])
    m, k = annot.shape
    blocks = np.asarray(blocks, dtype=np.int64)
    if blocks.shape != (m,):
        raise ValueError("block ids must give one block per SNP.")
    J = int(blocks.max()) + 1 if m else 0
    if J < 2 or set(np.unique(blocks).tolist()) != set(range(J)):
        raise ValueError("block ids must be the integers 0..J-1 with J >= 2.")
    names = _component_names(k, d.get("annotation_names"))

    def by_block(values):  # (J, ...) sums of per-SNP rows
        out = np.zeros((J,) + values.shape[1:])
        np.add.at(out, blocks, values)
        return out

    mass_b = by_block(annot)
    mass = annot.sum(axis=0)
    remain = mass[None, :] - mass_b
    if np.any(remain <= 0.0):
        jb, col = np.argwhere(remain <= 0.0)[0].tolist()
        raise ValueError(f"deleting block {jb} leaves annotation column {col} without mass.")
    weighted = {key: by_block(annot * (vec[key] ** 2 if key.startswith("score") else vec[key])[:, None])
                for key in vec}
    full_w = {key: v.sum(axis=0) for key, v in weighted.items()}
    cross_b = {}
    for key, p in pan.items():
        cb = np.zeros((J, k, k))
        for a in range(k):
            cb[:, a, :] = by_block(annot[:, a, None] * p)
        cross_b[key] = cb
    cross_full = {key: cb.sum(axis=0) for key, cb in cross_b.items()}
    out = []
    for jb in range(J):
        lhs, rhs, tr = _system(k, float(r), remain[jb],
                               full_w["score_x"] - weighted["score_x"][jb],
                               full_w["score_w"] - weighted["score_w"][jb],
                               full_w["norm_x"] - weighted["norm_x"][jb], full_w["norm_w"] - weighted["norm_w"][jb],
                               full_w["diag_nxe_x"] - weighted["diag_nxe_x"][jb],
                               full_w["diag_nxe_w"] - weighted["diag_nxe_w"][jb],
                               {key: cross_full[key] - cross_b[key][jb] for key in cross_b}, qn, qr, tn, tn2)
        out.append(GxENormalEquations(lhs, rhs, tr, names))
    return out


def fit(inputs: dict, blocks=None, *, max_condition: float = 1.0e12,
        allow_ill_conditioned: bool = False) -> GxEFitResult:
    """assemble_normal_equations + solve_normal_equations on `inputs` (its keyword
    arguments), with delete-one block jackknife SEs when `blocks` (block id per
    SNP) is given."""
    eq = assemble_normal_equations(**inputs)
    res = solve_normal_equations(eq, max_condition=max_condition, allow_ill_conditioned=allow_ill_conditioned)
    if blocks is None:
        return res
    reps = [solve_normal_equations(e, max_condition=max_condition, allow_ill_conditioned=allow_ill_conditioned)
            for e in deleted_block_equations(inputs, blocks)]
    cj = np.asarray([x.coefficients for x in reps])
    pj = np.asarray([x.proportions for x in reps])
    J = len(reps)

    def se(v):
        return np.sqrt((J - 1.0) / J * np.sum((v - v.mean(axis=0)) ** 2, axis=0))

    return replace(res, coefficient_standard_errors=se(cj), proportion_standard_errors=se(pj),
                   jackknife_coefficients=cj, jackknife_proportions=pj)


# --------------------------------------------------------------------------- design traces
def nxe_traces(env, C=None):
    """tr(P D P) and tr((P D P)^2) exactly, D = diag(e^2), P = I - Q Q' with
    Q = [1/sqrt(N), C] (C orthonormal and centred)
    (src/summit/ldscore/gxe_score.py:1213-1236)."""
    e = np.asarray(env, dtype=np.float64).ravel()
    n = e.size
    Q = np.full((n, 1), 1.0 / math.sqrt(n))
    if C is not None and np.asarray(C).size:
        Q = np.column_stack([Q, np.asarray(C, dtype=np.float64)])
    d = e * e
    QDQ = Q.T @ (d[:, None] * Q)
    sd, sd2 = float(d.sum()), float((d * d).sum())
    t1 = sd - float(np.trace(QDQ))
    t2 = sd2 - 2.0 * float(np.sum((d[:, None] * Q) ** 2)) + float(np.sum(QDQ * QDQ.T))
    tol = 2.0e-10 * max(1.0, sd, sd2)
    if t1 < -tol or t2 < -tol:
        raise RuntimeError("the NxE projection gave a negative trace.")
    return max(0.0, t1), max(0.0, t2)
