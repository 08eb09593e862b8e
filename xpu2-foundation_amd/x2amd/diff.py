"""Differentiable rendering through the device: a torch.autograd.Function whose forward renders one frame over the
C-ABI (x2dev_render_frame, x2dev_resolve_image) and whose backward runs the path-replay adjoint
(x2dev_adjoint_frame). The adjoint recomputes every path from its sample identity (pixel, sample index, frame seed)
instead of storing it, so the backward costs device time, not memory.

Parameters: one albedo texture (an H x W x 3 float32 tensor, texture 0 of the scene) and the roughness of a list of
materials. Torch only supplies device memory here; the library runs on its own HIP stream and every call into it
returns with that stream drained, so the caller synchronises torch's stream before each call (done below)."""
import torch

from . import abi
from .lane import GpuLane


class DiffRenderer:
    """A prepared lane over a scene whose texture 0 and whose materials `roughness_materials` are parameters."""

    def __init__(self, scene, config, roughness_materials, device=0, samples_per_chunk=4):
        if not torch.cuda.is_available():
            raise RuntimeError("DiffRenderer needs a GPU (the HIP device library)")
        self.scene, self.config = scene, config
        self.samples_per_chunk = samples_per_chunk
        self.roughness_materials = list(roughness_materials)
        self.materials = [abi.Material.from_buffer_copy(scene.materials[m]) for m in range(scene.material_count)]
        self.device = torch.device("cuda", device)
        self.lane = GpuLane(device)
        self.lane.prepare(scene)
        self.last_stats = None

    def close(self):
        self.lane.close()

    def _bind(self, texture, roughness):
        """Uploads the parameters the next frame (or adjoint) reads."""
        tex = texture.detach().to(self.device, torch.float32).contiguous()
        values = roughness.detach().cpu().tolist() if torch.is_tensor(roughness) else list(roughness)
        torch.cuda.synchronize(self.device)
        self.lane.set_texture(0, tex.data_ptr(), on_device=True)
        for k, m in enumerate(self.roughness_materials):
            mat = self.materials[m]
            mat.roughness = float(values[k])
            self.lane.set_material(m, mat)
        return tex

    def forward_image(self, texture, roughness):
        """Renders the frame; returns the image as an H x W x 3 float32 tensor on the device."""
        self._bind(texture, roughness)
        _, _, self.last_stats = self.lane.render_frame(self.config, self.samples_per_chunk)
        image = torch.empty((self.config.height, self.config.width, 3), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        self.lane.resolve_image(image.data_ptr())
        return image

    def adjoint(self, texture, roughness, grad_image):
        """dLoss/dtexture (H x W x 3) and dLoss/droughness (per listed material) for dLoss/dimage = grad_image."""
        tex = self._bind(texture, roughness)
        g = grad_image.detach().to(self.device, torch.float32).contiguous()
        grad_tex = torch.zeros_like(tex)
        grad_rough = torch.zeros(self.scene.material_count, dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        self.last_stats = self.lane.adjoint_frame(self.config, g.data_ptr(), grad_tex.data_ptr(), grad_rough.data_ptr(),
                                                  self.samples_per_chunk)
        return grad_tex, grad_rough[self.roughness_materials]

    def __call__(self, texture, roughness):
        return RenderFunction.apply(texture, roughness, self)


class RenderFunction(torch.autograd.Function):
    """image = render(texture, roughness); backward = the path-replay adjoint at the same sample identities."""

    @staticmethod
    def forward(ctx, texture, roughness, renderer):
        ctx.renderer = renderer
        ctx.save_for_backward(texture, roughness)
        return renderer.forward_image(texture, roughness)

    @staticmethod
    def backward(ctx, grad_image):
        texture, roughness = ctx.saved_tensors
        grad_tex, grad_rough = ctx.renderer.adjoint(texture, roughness, grad_image)
        return (grad_tex.to(texture.device, texture.dtype) if ctx.needs_input_grad[0] else None,
                grad_rough.to(roughness.device, roughness.dtype) if ctx.needs_input_grad[1] else None, None)
