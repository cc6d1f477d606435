"""Phase tables, local frames and general-C21 coefficient storage (host-side builders).

Follows ``solver/src/materials/material_field.py`` and
``solver/src/materials/hybrid_material.py``:

* a phase is ``{'kind': 'isotropic', 'E', 'nu'}``, ``{'kind': 'orthotropic',
  'E1', ..., 'G23'}`` or ``{'kind': 'anisotropic', 'C': 6x6}`` with an optional
  ``'frame'`` of ``'global'`` (default) or ``'local'``
  (``material_field.py:18-54``);
* a 6x6 engineering-Voigt stiffness (order 11, 22, 33, 23, 13, 12) is stored as
  its 21 upper-triangle entries, row-major (``PAIRS``, ``material_field.py:14``);
* ``orientation`` is a rotation array broadcastable to ``phase_id.shape + (3, 3)``
  whose columns are the local axes in global coordinates, or
  ``{'axis1': (..., 3), 'axis2': (..., 3)}`` (``material_field.py:71-102``); it
  is read only at voxels of local-frame anisotropic phases;
* ``build_material_field`` gives one row per voxel, (B, N, 21)
  (``material_field.py:117-165``); ``build_hybrid_material`` gives a pool with
  one row per phase plus one per local-frame anisotropic voxel and an int32
  (B, N) row index (``hybrid_material.py:14-70``).

Rotation is C'_ijkl = R_ia R_jb R_kc R_ld C_abcd on the full fourth-order
tensor, evaluated in FP64 on the host in chunks and copied to the device.
"""
from collections.abc import Mapping
from numbers import Integral

import numpy as np
import torch

PAIRS = tuple((i, j) for i in range(6) for j in range(i, 6))
ORTHOTROPIC_KEYS = ('E1', 'E2', 'E3', 'nu12', 'nu13', 'nu23', 'G12', 'G13', 'G23')
_ROWS = np.array([i for i, _ in PAIRS])
_COLS = np.array([j for _, j in PAIRS])
# Voigt index of the tensor index pair (i, j)
_VOIGT = np.array([[0, 5, 4], [5, 1, 3], [4, 3, 2]])


def isotropic_voigt(E, nu):
    """6x6 Voigt stiffness of an isotropic phase: lam = E nu/((1+nu)(1-2nu)), mu = E/(2(1+nu))."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2.0 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def orthotropic_voigt(E1, E2, E3, nu12, nu13, nu23, G12, G13, G23):
    """Inverse of the orthotropic compliance (S_12 = -nu12/E1, S_13 = -nu13/E1, S_23 = -nu23/E2)."""
    S = np.diag([1.0 / E1, 1.0 / E2, 1.0 / E3, 1.0 / G23, 1.0 / G13, 1.0 / G12])
    S[0, 1] = S[1, 0] = -nu12 / E1
    S[0, 2] = S[2, 0] = -nu13 / E1
    S[1, 2] = S[2, 1] = -nu23 / E2
    return np.linalg.inv(S)


def pack_c21(C):
    """(..., 6, 6) -> (..., 21) upper-triangle rows."""
    C = np.asarray(C, dtype=np.float64)
    return C[..., _ROWS, _COLS]


def unpack_c21(p):
    """(..., 21) -> symmetric (..., 6, 6)."""
    p = np.asarray(p, dtype=np.float64)
    C = np.zeros(p.shape[:-1] + (6, 6))
    C[..., _ROWS, _COLS] = p
    C[..., _COLS, _ROWS] = p
    return C


def material_stiffness(spec):
    """Validated symmetric positive-definite 6x6 stiffness of one phase (``material_field.py:18-54``)."""
    kind = spec.get('kind')
    if spec.get('frame', 'global') not in ('global', 'local'):
        raise ValueError("Material frame must be 'global' or 'local'")
    if kind == 'isotropic':
        E, nu = float(spec['E']), float(spec['nu'])
        if not (np.isfinite([E, nu]).all() and E > 0 and -1 < nu < 0.5):
            raise ValueError('Isotropic elasticity requires E>0 and -1<nu<0.5')
        C = isotropic_voigt(E, nu)
    elif kind == 'orthotropic':
        v = np.array([spec[k] for k in ORTHOTROPIC_KEYS], dtype=np.float64)
        if not np.isfinite(v).all() or (v[[0, 1, 2, 6, 7, 8]] <= 0).any():
            raise ValueError('Orthotropic moduli must be finite and positive')
        try:
            C = orthotropic_voigt(*v)
        except np.linalg.LinAlgError as exc:
            raise ValueError('Singular orthotropic compliance') from exc
    elif kind == 'anisotropic':
        C = np.asarray(spec['C'], dtype=np.float64)
    else:
        raise ValueError('kind must be isotropic, orthotropic or anisotropic')
    if C.shape != (6, 6) or not np.isfinite(C).all():
        raise ValueError('C must be a finite engineering-Voigt 6x6 matrix')
    if not np.allclose(C, C.T, rtol=0, atol=1e-12 * np.max(np.abs(C))):
        raise ValueError('Elastic stiffness must be symmetric')
    C = 0.5 * (C + C.T)
    try:
        np.linalg.cholesky(C)
    except np.linalg.LinAlgError as exc:
        raise ValueError('Elastic stiffness must be positive definite') from exc
    return C


def phase_table(materials):
    """Sorted int64 labels, their (P, 6, 6) stiffnesses and local-frame flags (``material_field.py:57-68``)."""
    if not isinstance(materials, Mapping) or not materials:
        raise ValueError('materials must be a nonempty mapping of integer labels')
    if any(not isinstance(k, Integral) or isinstance(k, bool) for k in materials):
        raise ValueError('Material labels must be integers')
    labels = np.array(sorted(materials), dtype=np.int64)
    specs = [materials[int(k)] for k in labels]
    tensors = np.stack([material_stiffness(s) for s in specs])
    local = np.array([s['kind'] != 'isotropic' and s.get('frame', 'global') == 'local' for s in specs])
    return labels, tensors, local


def orientation_frames(orientation, shape, flat_indices):
    """Rotation matrices (k, 3, 3) at the given flat voxel indices (``material_field.py:71-102``)."""
    where = np.unravel_index(np.asarray(flat_indices), shape)

    def pick(values, tail):
        try:
            view = np.broadcast_to(np.asarray(values, dtype=np.float64), tuple(shape) + tail)
        except ValueError as exc:
            raise ValueError('Orientation shape is not broadcastable to phase_id') from exc
        out = np.array(view[where], copy=True)
        if not np.isfinite(out).all():
            raise ValueError('Active local-frame orientations must be finite')
        return out

    if isinstance(orientation, Mapping):
        a, b = pick(orientation['axis1'], (3,)), pick(orientation['axis2'], (3,))
        na, nb = np.linalg.norm(a, axis=-1), np.linalg.norm(b, axis=-1)
        if (na <= 0).any() or (nb <= 0).any():
            raise ValueError('Orientation axes must be nonzero')
        a = a / na[:, None]
        b = b / nb[:, None]
        c = np.cross(a, b)
        nc = np.linalg.norm(c, axis=-1)
        if (nc <= 1e-12).any():
            raise ValueError('Orientation axes must not be collinear')
        c = c / nc[:, None]
        # columns: axis1, the part of axis2 orthogonal to it, and their normal
        return np.stack([a, np.cross(c, a), c], axis=-1)
    R = pick(orientation, (3, 3))
    if not np.allclose(np.swapaxes(R, -1, -2) @ R, np.eye(3), rtol=0, atol=1e-10):
        raise ValueError('Orientation must be an orthonormal rotation matrix')
    if not np.allclose(np.linalg.det(R), 1, rtol=0, atol=1e-10):
        raise ValueError('Orientation must be right-handed (det=+1)')
    return R


def voigt_to_tensor(C):
    """(6, 6) engineering-Voigt stiffness -> (3, 3, 3, 3) tensor C_ijkl."""
    return np.asarray(C, dtype=np.float64)[_VOIGT[:, :, None, None], _VOIGT[None, None, :, :]]


def rotate_c21(C, R):
    """Packed rows (k, 21) of R C R^T for one 6x6 stiffness C and k rotations R (k, 3, 3)."""
    T = voigt_to_tensor(C)
    R = np.asarray(R, dtype=np.float64)
    T = np.einsum('nia,abcd->nibcd', R, T)
    T = np.einsum('njb,nibcd->nijcd', R, T)
    T = np.einsum('nkc,nijcd->nijkd', R, T)
    T = np.einsum('nld,nijkd->nijkl', R, T)
    ij = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    out = np.empty((R.shape[0], 21))
    for k, (I, J) in enumerate(PAIRS):
        (i, j), (kk, ll) = ij[I], ij[J]
        out[:, k] = T[:, i, j, kk, ll]
    return out


def _check_inputs(phase_id, chunk_size, dtype):
    phase = np.asarray(phase_id)
    if phase.ndim not in (3, 4) or 0 in phase.shape or phase.dtype.kind not in 'iu':
        raise ValueError('phase_id must be a nonempty integer (Nz,Ny,Nx) or (B,Nz,Ny,Nx) array')
    if not isinstance(chunk_size, Integral) or isinstance(chunk_size, bool) or chunk_size < 1:
        raise ValueError('chunk_size must be a positive integer')
    dtype = torch.float64 if dtype is None else dtype
    if dtype not in (torch.float32, torch.float64):
        raise ValueError('Material dtype must be float32 or float64')
    return phase, dtype


def _phase_rows(phase, materials, orientation):
    labels, tensors, local = phase_table(materials)
    used = np.unique(phase)
    known = np.isin(used, labels)
    if not known.all():
        raise ValueError(f'Undefined material labels: {used[~known].tolist()}')
    if orientation is None and local[np.searchsorted(labels, used)].any():
        raise ValueError('Active local-frame anisotropic phases require orientation')
    return labels, tensors, local, np.searchsorted(labels, phase.reshape(-1))


def build_material_field(phase_id, materials, *, orientation=None, device='cuda', dtype=None, chunk_size=65536):
    """One C21 row per voxel, (B, N, 21) on ``device`` (``material_field.py:117-165``)."""
    phase, dtype = _check_inputs(phase_id, chunk_size, dtype)
    labels, tensors, local, rows = _phase_rows(phase, materials, orientation)
    table = pack_c21(tensors)
    field = torch.empty((rows.size, 21), device=device, dtype=dtype)
    for start in range(0, rows.size, chunk_size):
        stop = min(start + chunk_size, rows.size)
        block = table[rows[start:stop]]
        for p in np.flatnonzero(local):
            sel = np.flatnonzero(rows[start:stop] == p)
            if sel.size:
                block[sel] = rotate_c21(tensors[p], orientation_frames(orientation, phase.shape, start + sel))
        field[start:stop].copy_(torch.from_numpy(block))
    B = phase.shape[0] if phase.ndim == 4 else 1
    return field.reshape(B, -1, 21)


def build_hybrid_material(phase_id, materials, *, orientation=None, device='cuda', dtype=None, chunk_size=65536):
    """``(pool, index, metadata)`` (``hybrid_material.py:14-70``).

    pool (P + M, 21): one row per phase (sorted labels) followed by one rotated
    row per voxel of a local-frame anisotropic phase (M of them, phase by
    phase in voxel order); index int32 (B, N) is the row of each voxel.
    """
    phase, dtype = _check_inputs(phase_id, chunk_size, dtype)
    labels, tensors, local, rows = _phase_rows(phase, materials, orientation)
    count = int(np.count_nonzero(local[rows]))
    total = len(labels) + count
    if total > np.iinfo(np.int32).max:
        raise ValueError('Hybrid coefficient pool exceeds int32 indexing')
    index = rows.astype(np.int32)
    pool = torch.empty((total, 21), device=device, dtype=dtype)
    pool[:len(labels)].copy_(torch.from_numpy(pack_c21(tensors)))
    offset = len(labels)
    for p in np.flatnonzero(local):
        sel = np.flatnonzero(rows == p)
        if not sel.size:
            continue
        index[sel] = np.arange(offset, offset + sel.size, dtype=np.int32)
        for start in range(0, sel.size, chunk_size):
            stop = min(start + chunk_size, sel.size)
            block = rotate_c21(tensors[p], orientation_frames(orientation, phase.shape, sel[start:stop]))
            pool[offset + start:offset + stop].copy_(torch.from_numpy(block))
        offset += sel.size
    B = phase.shape[0] if phase.ndim == 4 else 1
    index_dev = torch.from_numpy(index.reshape(B, -1)).to(device)
    itemsize = pool.element_size()
    metadata = dict(labels=labels.tolist(), n_elements=int(phase.size), n_local_voxels=count,
                    n_shared_voxels=int(phase.size - count), pool_rows=total,
                    coefficient_bytes=pool.numel() * itemsize, index_bytes=index_dev.numel() * 4,
                    dense_bytes=int(phase.size) * 21 * itemsize)
    return pool, index_dev, metadata
