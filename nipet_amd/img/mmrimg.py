"""Image layout conversions and the reconstruction mask.

* ``convert2e7`` / ``convert2dev`` -- ``niftypet/nipet/img/mmrimg.py:28-67``:
  Siemens/E7 image (z=127, y=344, x=344) <-> device image (y=320, x=320,
  z=127): crop/pad a 12-voxel transaxial margin and move z to the fastest axis
  (the projectors read whole z-columns).
* ``get_cylinder`` -- ``niftypet/nipet/img/mmrimg.py:208-227``: a uniform
  transaxial disc, used as the OSEM mask (radius 29 cm by default).
"""
import math

import numpy as np

__all__ = ['convert2e7', 'convert2dev', 'get_cylinder', 'cylinder_disc']


def convert2e7(img, Cnt):
    """Device image (320, 320, nz) -> Siemens image (nz, 344, 344)."""
    if getattr(img, 'is_cuda', False):
        img = img.detach().cpu().numpy()
    img = np.asarray(img)
    margin = (Cnt['SO_IMX'] - Cnt['SZ_IMX']) // 2
    nvz = img.shape[2]
    out = np.zeros((nvz, Cnt['SO_IMY'], Cnt['SO_IMX']), dtype=np.float32)
    out[:, margin:margin + Cnt['SZ_IMY'], margin:margin + Cnt['SZ_IMX']] = \
        np.transpose(img, (2, 0, 1))
    return out


def convert2dev(im, Cnt):
    """Siemens image (nz, 344, 344) -> device image (320, 320, nz), contiguous."""
    im = np.asarray(im)
    if im.ndim != 3 or im.shape[1] != Cnt['SO_IMY'] or im.shape[2] != Cnt['SO_IMX']:
        raise ValueError('input image array is not of the Siemens shape (nz, 344, 344)')
    margin = (Cnt['SO_IMX'] - Cnt['SZ_IMX']) // 2
    sub = im[:, margin:margin + Cnt['SZ_IMY'], margin:margin + Cnt['SZ_IMX']]
    return np.ascontiguousarray(np.transpose(sub, (1, 2, 0)), dtype=np.float32)


def cylinder_disc(Cnt, rad=25, xo=0, yo=0, unival=1):
    """Transaxial (344, 344) slice of ``get_cylinder``.

    The disc is rasterised by sweeping half-angles t in [0, pi) in steps of
    pi/720 and filling, for each t, the column u(x=xo+rad cos t) over the rows
    covered by y in [-y+2yo, y) sampled at half a voxel, with
    u = 172 + floor(x/vox), v = 172 - ceil(y/vox) (the same sampling as the
    reference, ``img/mmrimg.py:208-227``, so the mask voxels agree)."""
    vox = Cnt['SO_VXY']
    nx, ny = Cnt['SO_IMX'], Cnt['SO_IMY']
    disc = np.zeros((nx, ny), dtype=np.float32)
    for t in np.arange(0, math.pi, math.pi / (2*360)):
        x = xo + rad * math.cos(t)
        y = yo + rad * math.sin(t)
        yf = np.arange(-y + 2*yo, y, vox / 2)
        v = np.int32(.5 * nx - np.ceil(yf / vox))
        u = np.int32(.5 * ny + np.floor(x / vox))
        disc[v, u] = unival
    return disc


def get_cylinder(Cnt, rad=25, xo=0, yo=0, unival=1, gpu_dim=False):
    """Uniform cylinder of value ``unival``, transaxial radius ``rad`` [cm],
    centre (xo, yo) [cm], in the Siemens shape (or device shape if
    ``gpu_dim``): ``cylinder_disc`` repeated over the planes."""
    disc = cylinder_disc(Cnt, rad, xo, yo, unival)
    nvz = Cnt.get('rSO_IMZ', Cnt['SO_IMZ'])
    img = np.repeat(disc[None], nvz, axis=0)
    if gpu_dim:
        img = convert2dev(img, Cnt)
    return img
