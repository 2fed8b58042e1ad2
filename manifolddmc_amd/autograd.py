"""Differentiable dual marching cubes: torch.autograd.Function over the C ABI.

The forward is mdmc_extract with RECORD_ADJOINT (dual vertices as edge means, no
move-to-surface iteration, no vertex-changing post-process); the backward is
mdmc_backward, a HIP adjoint that maps dL/d(vertex position) onto the input field
with the topology of the forward frozen.  Both run on torch's current stream.
"""
import torch

from . import PAD1, RECORD_ADJOINT, DualMarchingCubes, MdmcError

_default = {}


def _context(device_index):
    d = _default.get(device_index)
    if d is None:
        d = _default[device_index] = DualMarchingCubes(device=device_index)
    return d


class DMCFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, iso, origin, delta, pad, refinement, dmc):
        if field.dtype != torch.float32 or not field.is_cuda or field.dim() != 3:
            raise ValueError("field must be a 3-D float32 tensor on a GPU")
        field = field.contiguous()
        dmc.set_stream(torch.cuda.current_stream(field.device).cuda_stream)
        nz, ny, nx = field.shape
        flags = RECORD_ADJOINT | (PAD1 if pad else 0)
        dmc.extract_raw(field.data_ptr(), (nx, ny, nz), origin, delta, float(iso), flags, refinement, True)
        V, H, F = dmc.info()
        pos = torch.empty((V, 3), dtype=torch.float32, device=field.device)
        he = torch.empty((H, 4), dtype=torch.int32, device=field.device)
        fh = torch.empty((F,), dtype=torch.int32, device=field.device)
        if V:
            dmc.copy_device(positions=pos.data_ptr(), halfedges=he.data_ptr() if H else 0, face_halfedge=fh.data_ptr() if F else 0)
        # quads from the halfedge cycles: first halfedge of each face, then three next steps
        quads = torch.empty((F, 4), dtype=torch.int64, device=field.device)
        if F:
            h = fh.long()
            for i in range(4):
                quads[:, i] = he[h, 0].long()
                h = he[h, 2].long()
        ctx.dmc = dmc
        ctx.generation = dmc.generation
        ctx.shape = field.shape
        ctx.device = field.device
        ctx.mark_non_differentiable(quads)
        return pos, quads

    @staticmethod
    def backward(ctx, gpos, gquads):
        dmc = ctx.dmc
        if dmc.generation != ctx.generation:
            raise MdmcError(6, "the DMC context ran another extraction after this forward; backward needs the latest one")
        dfield = torch.empty(ctx.shape, dtype=torch.float32, device=ctx.device)
        g = gpos.contiguous() if gpos is not None else torch.zeros((0, 3), dtype=torch.float32, device=ctx.device)
        dmc.set_stream(torch.cuda.current_stream(ctx.device).cuda_stream)
        dmc.backward(g.data_ptr() if g.numel() else 0, dfield.data_ptr(), False)
        return dfield, None, None, None, None, None, None


def dual_marching_cubes(field, iso, origin=(0.0, 0.0, 0.0), delta=(1.0, 1.0, 1.0), pad=True, refinement=0, dmc=None):
    """Differentiable DMC of a 3-D float32 CUDA tensor [z][y][x]: returns (positions [V,3]
    float32 with grad, quads [F,4] int64 vertex ids)."""
    if dmc is None:
        dmc = _context(field.device.index if field.device.index is not None else torch.cuda.current_device())
    return DMCFunction.apply(field, iso, tuple(origin), tuple(delta), bool(pad), int(refinement), dmc)
