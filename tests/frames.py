"""Builds the C frame structs of include/rmh.h over numpy planes (host memory)."""
import numpy as np

from conftest import load_package

abi = load_package().abi


def dbk_frame_struct(planes, edges, bd, qps, beta=(0, 0, 0), tc=(0, 0, 0)):
    """planes: [Y, Cb, Cr] C-contiguous arrays (modified in place by a run);
    edges: uint8[gh, gw, 4]. Returns (struct, keepalive)."""
    edges = np.ascontiguousarray(edges, dtype=np.uint8)
    ps = [abi.plane_from_array(p, bd) for p in planes]
    f = abi.dbk_frame(ps, edges.ctypes.data, edges.shape[1], edges.shape[0], qps, beta, tc)
    return f, (planes, edges)


def mc_frame_struct(refs, dsts, bd):
    f = abi.mc_frame([abi.plane_from_array(p, bd) for p in refs], [abi.plane_from_array(p, bd) for p in dsts])
    return f, (refs, dsts)


def mc_bi_frame_struct(refs0, refs1, dsts, bd):
    f = abi.McBiFrame((abi.Plane * 3)(*[abi.plane_from_array(p, bd) for p in refs0]),
                      (abi.Plane * 3)(*[abi.plane_from_array(p, bd) for p in refs1]),
                      (abi.Plane * 3)(*[abi.plane_from_array(p, bd) for p in dsts]))
    return f, (refs0, refs1, dsts)


def recon_frame_struct(planes, bd):
    """planes: [Y, Cb, Cr] arrays that a recon run adds residuals into."""
    f = abi.ReconFrame((abi.Plane * 3)(*[abi.plane_from_array(p, bd) for p in planes]))
    return f, planes
