"""The textured scene and the path-replay adjoint on the device, through the C-ABI.

- forward parity of the textured Cornell box against the oracle (counts exact, radiance within 1e-3);
- the device adjoint against the oracle's FP64 adjoint at the same sample identities;
- the device adjoint against central differences (h = 1e-2: the step the FP32 mip levels resolve) of the device's
  own renders on 64 random texels of the
  differentiable Cornell box at its full size (512 x 512, 32 spp, depth 8, 512 x 512 texture, L2 loss against a
  target rendered from parameters perturbed by N(0, 0.1)): agreement within 1e-2 relative;
- the roughness gradients against differences taken with the sampling held fixed (the estimator the adjoint
  differentiates; the oracle's evalRoughness hook), within 1e-2 relative;
- the torch.autograd.Function over the same entry points.
Device buffers come from torch (device memory only)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from helpers import mean_relative_difference
from x2amd import abi, scenes
from x2amd.lane import GpuLane

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def image_of(lane, config):
    sums, counts, st = lane.render_frame(config)
    assert (counts == config.spp).all() and st.errors == 0
    return sums[:, :3] / counts[:, None].astype(np.float64)


def device_adjoint(lane, scene, config, d_image, texture_floats):
    g = torch.from_numpy(np.ascontiguousarray(d_image, np.float32).reshape(-1)).to(DEV)
    gt = torch.zeros(max(1, texture_floats), dtype=torch.float32, device=DEV)
    gr = torch.zeros(scene.material_count, dtype=torch.float32, device=DEV)
    torch.cuda.synchronize()
    st = lane.adjoint_frame(config, g.data_ptr(), gt.data_ptr(), gr.data_ptr())
    assert st.errors == 0
    return gt.cpu().numpy().astype(np.float64)[:texture_floats], gr.cpu().numpy().astype(np.float64), st


@pytest.fixture(scope="module")
def small():
    scene, config, params = scenes.cornell_textured(width=128, height=128, spp=8, max_depth=8)
    lane = GpuLane(0)
    lane.prepare(scene)
    yield scene, config, params, lane
    lane.close()


def test_textured_forward_parity(small):
    scene, config, params, lane = small
    sums, counts, st = lane.render_frame(config)
    ref_sums, ref_counts, ref_k = oracle.OracleScene(scene).render(config)
    assert (counts == config.spp).all() and (ref_counts == config.spp).all()
    rel = mean_relative_difference(sums, counts, ref_sums, ref_counts)
    print(f"textured cornell 128x128x8: rel {rel:.3e}, traces {st.rays_trace} vs {ref_k[0]}")
    assert rel <= 1e-3
    assert st.rays_trace == ref_k[0] and st.rays_shadow == ref_k[1]


def test_adjoint_matches_oracle_fp64(small):
    scene, config, params, lane = small
    rng = np.random.default_rng(8)
    d_image = rng.normal(size=(config.width * config.height, 3)).astype(np.float32)
    n = params["texture"].size
    gt, gr, st = device_adjoint(lane, scene, config, d_image, n)
    ot, orr, ok = oracle.OracleScene(scene).adjoint(config, d_image, n, scene.material_count)
    assert ok[2] == 0
    assert st.rays_trace == ok[0] and st.rays_shadow == ok[1], "the replay traces the forward paths"
    err = np.linalg.norm(gt - ot) / np.linalg.norm(ot)
    print(f"texel gradient vs FP64 oracle: relative L2 {err:.3e}; roughness {gr[1:5]} vs {orr[1:5]}")
    assert err <= 1e-4
    assert np.all(np.abs(gt - ot) <= 1e-3 * np.abs(ot) + 1e-5 * np.abs(ot).max())
    assert np.allclose(gr, orr, rtol=1e-4, atol=1e-6 * np.abs(orr).max())


def test_roughness_adjoint_matches_detached_differences(small):
    """Differences over the lobe's value with sampling held (the oracle's hook) at the same sample identities."""
    scene, config, params, lane = small
    rng = np.random.default_rng(9)
    d_image = rng.normal(size=(config.width * config.height, 3)).astype(np.float32)
    _, gr, _ = device_adjoint(lane, scene, config, d_image, params["texture"].size)
    o = oracle.OracleScene(scene)
    h = 1e-3

    def loss():
        sums, counts, _ = o.render(config)
        return float(np.sum(sums[:, :3] / counts[:, None] * d_image))

    for m, r in zip(params["roughness_materials"], params["roughness"]):
        o.set_eval_roughness(m, float(r) + h)
        lp = loss()
        o.set_eval_roughness(m, float(r) - h)
        lm = loss()
        o.set_eval_roughness(m, -1.0)
        fd = (lp - lm) / (2 * h)
        print(f"material {m}: device adjoint {gr[m]:.6e}, detached difference {fd:.6e}")
        assert abs(gr[m] - fd) <= 1e-2 * abs(fd), (m, gr[m], fd)


@pytest.fixture(scope="module")
def full():
    """The differentiable Cornell box at full size and its target."""
    scene, config, params = scenes.cornell_textured()
    lane = GpuLane(0)
    lane.prepare(scene)
    rng = np.random.default_rng(1234)
    tex_t = np.clip(params["texture"] + rng.normal(0, 0.1, params["texture"].shape), 0.01, 1.0).astype(np.float32)
    rough_t = np.clip(params["roughness"] + rng.normal(0, 0.1, 4), 0.05, 0.95)
    lane.set_texture(0, tex_t)
    for m, r in zip(params["roughness_materials"], rough_t):
        mat = abi.Material.from_buffer_copy(scene.materials[m])
        mat.roughness = float(r)
        lane.set_material(m, mat)
    target = image_of(lane, config)
    lane.set_texture(0, params["texture"])
    for m in params["roughness_materials"]:
        lane.set_material(m, scene.materials[m])
    yield scene, config, params, lane, target
    lane.close()


def test_texel_gradients_match_central_differences(full):
    """Central differences of the loss in 64 random texels against the adjoint. The step is 1e-2: the forward frame
    reads the texture's FP32 mip levels, where a texel's step moves a level-k texel by h / 4^k, and at h = 1e-3 the
    coarse levels (the 1 x 1 level moves by h / 262144) round most of the move away, so the differences lose the
    gradient that flows through them (the oracle shows the same: 7e-3 relative at h = 1e-3, 1.3e-3 at 1e-2). The
    estimator is a polynomial in the texels, so the larger step adds no truncation error worth the name."""
    scene, config, params, lane, target = full
    image = image_of(lane, config)
    loss0 = float(np.sum((image - target) ** 2))
    d_image = (2 * (image - target)).astype(np.float32)
    tex = params["texture"]
    gt, gr, st = device_adjoint(lane, scene, config, d_image, tex.size)
    rng = np.random.default_rng(64)
    picks = rng.choice(tex.size, 64, replace=False)
    h = 1e-2
    fd = []
    flat = tex.reshape(-1).copy()
    for i in picks:
        flat[i] = tex.reshape(-1)[i] + h
        lane.set_texture(0, flat.reshape(tex.shape))
        lp = float(np.sum((image_of(lane, config) - target) ** 2))
        flat[i] = tex.reshape(-1)[i] - h
        lane.set_texture(0, flat.reshape(tex.shape))
        lm = float(np.sum((image_of(lane, config) - target) ** 2))
        flat[i] = tex.reshape(-1)[i]
        fd.append((lp - lm) / (2 * h))
    lane.set_texture(0, tex)
    fd, adj = np.array(fd), gt[picks]
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-12)
    big = np.abs(fd) >= 1e-2 * np.abs(fd).max()
    print(f"loss {loss0:.6e}; 64 texels: median rel {np.median(rel):.3e}, max rel (|fd| >= 1% of max) "
          f"{rel[big].max():.3e}, vector rel {np.linalg.norm(adj - fd) / np.linalg.norm(fd):.3e}")
    assert np.linalg.norm(adj - fd) <= 1e-2 * np.linalg.norm(fd)


def test_texel_block_gradients_match_central_differences(full):
    """Central differences over 16 random aligned 16 x 16 blocks of texels (every texel of the block moved by h)
    against the sum of the adjoint's texel gradients over the block. A block's move shifts each level above by at
    least h / 1024, which the FP32 levels resolve, so these differences carry the gradient that flows through the
    coarse levels too (the single texels' above cannot see it: the adjoint gives each texel 1 / 4^k of a level-k
    texel's gradient, some 3e-5 here from the coarsest levels that diffuse paths read)."""
    scene, config, params, lane, target = full
    image = image_of(lane, config)
    d_image = (2 * (image - target)).astype(np.float32)
    tex = params["texture"]
    gt, _, _ = device_adjoint(lane, scene, config, d_image, tex.size)
    g = gt.reshape(tex.shape)
    rng = np.random.default_rng(65)
    B, h = 16, 1e-2
    blocks = [(int(y), int(x), int(c)) for y, x, c in zip(rng.integers(0, tex.shape[0] // B, 16),
                                                         rng.integers(0, tex.shape[1] // B, 16), rng.integers(0, 3, 16))]
    fd, adj = [], []
    for by, bx, c in blocks:
        sl = (slice(by * B, (by + 1) * B), slice(bx * B, (bx + 1) * B), c)
        t = tex.copy()
        t[sl] += h
        lane.set_texture(0, t)
        lp = float(np.sum((image_of(lane, config) - target) ** 2))
        t[sl] -= 2 * h
        lane.set_texture(0, t)
        lm = float(np.sum((image_of(lane, config) - target) ** 2))
        fd.append((lp - lm) / (2 * h))
        adj.append(float(g[sl].sum()))
    lane.set_texture(0, tex)
    fd, adj = np.array(fd), np.array(adj)
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-12)
    print(f"16 blocks of 16x16: median rel {np.median(rel):.3e}, max rel {rel.max():.3e}, "
          f"vector rel {np.linalg.norm(adj - fd) / np.linalg.norm(fd):.3e}")
    assert np.linalg.norm(adj - fd) <= 1e-2 * np.linalg.norm(fd)
    assert np.all(np.abs(adj - fd) <= 1e-2 * np.abs(fd) + 1e-3 * np.abs(fd).max())


def test_roughness_gradients_at_full_size_over_a_window(full):
    """The roughness gradients at the differentiable Cornell box's full size (512x512, 32 spp, depth 8, RR off) for the
    loss restricted to a 64x64 window (the image adjoint 2 (image - target) there, zero elsewhere): the device's
    adjoint over the whole frame against the oracle's differences with sampling held fixed (its evalRoughness hook),
    rendered over the window alone at the same sample identities, h = 1e-3."""
    scene, config, params, lane, target = full
    image = image_of(lane, config)
    vp = [224, 288, 288, 352]
    y, x = np.divmod(np.arange(config.width * config.height), config.width)
    inside = (x >= vp[0]) & (x < vp[2]) & (y >= vp[1]) & (y < vp[3])
    d_image = np.zeros((config.width * config.height, 3), np.float32)
    d_image[inside] = 2 * (image[inside] - target[inside])
    _, gr, _ = device_adjoint(lane, scene, config, d_image, params["texture"].size)
    window = abi.Config.from_buffer_copy(config)
    window.viewport[:] = vp
    o = oracle.OracleScene(scene)
    h = 1e-3

    def loss():
        sums, counts, k = o.render(window)
        assert k[2] == 0 and (counts[inside] == config.spp).all()
        return float(np.sum(sums[inside, :3] / counts[inside, None] * d_image[inside]))

    fds = []
    for m, r in zip(params["roughness_materials"], params["roughness"]):
        o.set_eval_roughness(m, float(r) + h)
        lp = loss()
        o.set_eval_roughness(m, float(r) - h)
        lm = loss()
        o.set_eval_roughness(m, -1.0)
        fds.append((lp - lm) / (2 * h))
    scale = max(abs(f) for f in fds)
    for m, fd in zip(params["roughness_materials"], fds):
        print(f"material {m}: device adjoint {gr[m]:.6e}, detached difference {fd:.6e}")
        assert abs(gr[m] - fd) <= 1e-2 * abs(fd) + 1e-3 * scale, (m, gr[m], fd)


def test_autograd_function_steps_downhill(small):
    from x2amd.diff import DiffRenderer
    scene, config, params, _ = small
    r = DiffRenderer(scene, config, params["roughness_materials"])
    try:
        tex = torch.tensor(params["texture"], device=DEV, requires_grad=True)
        rough = torch.tensor(params["roughness"], device=DEV, requires_grad=True)
        with torch.no_grad():
            target = r(torch.clamp(tex + 0.1 * torch.randn_like(tex), 0.01, 1.0), rough + 0.05)
        image = r(tex, rough)
        loss = ((image - target) ** 2).sum()
        loss.backward()
        assert tex.grad is not None and rough.grad is not None
        assert torch.isfinite(tex.grad).all() and torch.isfinite(rough.grad).all()
        # the Function's backward is the adjoint entry point with dLoss/dimage = 2 (image - target)
        gt, gr = r.adjoint(tex, rough, 2 * (image.detach() - target))
        assert torch.allclose(tex.grad, gt, rtol=1e-4, atol=1e-6 * float(gt.abs().max()))
        assert torch.allclose(rough.grad, gr, rtol=1e-4)
        with torch.no_grad():
            step = 0.5 * float(loss) / float((tex.grad ** 2).sum())  # a small fraction of the linearised distance
            after = ((r(tex - 0.1 * step * tex.grad, rough) - target) ** 2).sum()
        print(f"loss {float(loss):.6e} -> {float(after):.6e} after one texture step")
        assert float(after) < float(loss)
    finally:
        r.close()


def test_adjoint_argument_errors(small):
    scene, config, params, lane = small
    g = torch.zeros(config.width * config.height * 3, device=DEV)
    gr = torch.zeros(scene.material_count, device=DEV)
    with pytest.raises(RuntimeError):
        lane.adjoint_frame(config, g.data_ptr(), 0, gr.data_ptr())  # texture gradient missing
    deep = scenes.cornell_textured(width=128, height=128, spp=8, max_depth=40)[1]
    gt = torch.zeros(params["texture"].size, device=DEV)
    with pytest.raises(RuntimeError):
        lane.adjoint_frame(deep, g.data_ptr(), gt.data_ptr(), gr.data_ptr())
    with pytest.raises(RuntimeError):
        lane.set_texture(3, params["texture"])  # no such texture


def test_diff_renderer_rejects_wrong_shapes(small):
    """The texture is copied from the tensor's memory at texture 0's size and the texel gradients are written at the
    scene's sizes, so a tensor of another shape (or an image gradient of another size) is refused before the device
    reads or writes it."""
    from x2amd.diff import DiffRenderer
    scene, config, params, _ = small
    r = DiffRenderer(scene, config, params["roughness_materials"])
    try:
        tex = torch.tensor(params["texture"], device=DEV)
        rough = torch.tensor(params["roughness"], device=DEV)
        with pytest.raises(ValueError):
            r.forward_image(tex[: tex.shape[0] // 2], rough)
        with pytest.raises(ValueError):
            r.forward_image(tex, rough[:-1])
        with pytest.raises(ValueError):
            r.adjoint(tex, rough, torch.zeros(config.height // 2, config.width, 3, device=DEV))
        gt, gr = r.adjoint(tex, rough, torch.zeros(config.height, config.width, 3, device=DEV))
        assert tuple(gt.shape) == tuple(params["texture"].shape) and gr.shape[0] == len(params["roughness"])
    finally:
        r.close()
