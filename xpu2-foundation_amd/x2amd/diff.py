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
        if not scene.texture_arrays:
            raise ValueError("DiffRenderer needs a scene with a texture (texture 0 is the parameter)")
        sizes = [(t.shape[0], t.shape[1]) for t in scene.texture_arrays]
        self.texture_shape = (sizes[0][0], sizes[0][1], 3)
        # the adjoint's texel gradients cover every texture of the scene, concatenated in order
        self.grad_texels = sum(h * w * 3 for h, w in sizes)
        for m in self.roughness_materials:
            if not 0 <= m < scene.material_count:
                raise ValueError(f"roughness material {m} is not a material of the scene")
        self.device = torch.device("cuda", device)
        self.lane = GpuLane(device)
        self.lane.prepare(scene)
        self.last_stats = None

    def close(self):
        self.lane.close()

    def _bind(self, texture, roughness):
        """Uploads the parameters the next frame (or adjoint) reads. The device copies texture 0's texels from the
        tensor's memory, so its shape must be texture 0's."""
        if tuple(texture.shape) != self.texture_shape:
            raise ValueError(f"texture of shape {tuple(texture.shape)}; texture 0 of the scene is {self.texture_shape}")
        tex = texture.detach().to(self.device, torch.float32).contiguous()
        values = roughness.detach().cpu().reshape(-1).tolist() if torch.is_tensor(roughness) else list(roughness)
        if len(values) != len(self.roughness_materials):
            raise ValueError(f"{len(values)} roughness values for {len(self.roughness_materials)} materials")
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
        self._bind(texture, roughness)
        image_shape = (self.config.height, self.config.width, 3)
        if tuple(grad_image.shape) != image_shape:
            raise ValueError(f"image gradient of shape {tuple(grad_image.shape)}; the image is {image_shape}")
        g = grad_image.detach().to(self.device, torch.float32).contiguous()
        # sized from the scene's textures (the kernel writes every texture's texel gradients), texture 0 returned
        grad_all = torch.zeros(self.grad_texels, dtype=torch.float32, device=self.device)
        grad_rough = torch.zeros(self.scene.material_count, dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        self.last_stats = self.lane.adjoint_frame(self.config, g.data_ptr(), grad_all.data_ptr(), grad_rough.data_ptr(),
                                                  self.samples_per_chunk)
        n0 = self.texture_shape[0] * self.texture_shape[1] * 3
        return grad_all[:n0].view(self.texture_shape), grad_rough[self.roughness_materials]

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
