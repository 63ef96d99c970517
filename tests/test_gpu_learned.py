"""Learned denoiser (BF16 MFMA kernels) against a plain PyTorch FP32 reference.

The reference is F.conv2d in FP32 with the same weights; with
`emulate_bf16` it rounds the packed input, the weights and every hidden
activation to BF16 at the points where the kernels store BF16, so the only
remaining difference is the FP32 accumulation order (tolerance 2e-3 in log
space).  Gradients: the kernels carry BF16 activation gradients, autograd
carries FP32 ones, so per-layer gradients agree to a few percent (relative
L2 <= 5e-2, cosine >= 0.998).
"""
import numpy as np
import pytest

from mcrt_amd import camera as cam_mod, configs, learned

pytestmark = pytest.mark.gpu


def torch_reference(w, noisy, albedo, gbuf, emulate_bf16=True):
    import torch
    import torch.nn.functional as F

    bf = (lambda t: t.to(torch.bfloat16).float()) if emulate_bf16 else (lambda t: t)
    L = torch.log1p(torch.clamp(noisy[..., :3], min=0.0))
    d = torch.clamp(gbuf[..., 3], 0.0, 1e4)
    x = torch.cat([L, albedo[..., :3], gbuf[..., :3], (torch.log1p(d) * (1.0 / 9.21044))[..., None]], -1)
    x = bf(x).permute(0, 3, 1, 2)
    out = None
    for l, (W, b) in enumerate(w):
        y = F.conv2d(x, bf(W), b, padding=1)
        if l < 5:
            x = bf(F.relu(y))
        else:
            out = L.permute(0, 3, 1, 2) + y
    return out.permute(0, 2, 3, 1)


@pytest.fixture(scope="module")
def batch(ctx):
    import torch

    dev = torch.device("cuda:0")
    B, S = 2, 48
    noisy = torch.empty((B, S, S, 4), device=dev)
    albedo = torch.empty_like(noisy)
    gbuf = torch.empty_like(noisy)
    ref = torch.empty_like(noisy)
    tmp = torch.empty((S, S, 4), device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    for i in range(B):
        loaded, sb = learned.variant_scene(77 + i, "c1" if i == 0 else "c2")
        cam = cam_mod.camera_for(loaded, S, S)
        ctx.upload_scene(sb)
        ctx.render_aux_dev(cam, S, S, 4, 0, 8, None, noisy[i].data_ptr(), gbuf[i].data_ptr(),
                           albedo[i].data_ptr(), s)
        learned.reference_render(ctx, cam, S, S, 4, 64, 8, noisy[i], ref[i], tmp, s)
    torch.cuda.synchronize()
    return noisy, albedo, gbuf, ref


def _weights_torch(dn, requires_grad=False):
    import torch

    out = []
    for W, b in learned.split_weights(dn.get_weights()):
        tw = torch.tensor(W, device="cuda:0", requires_grad=requires_grad)
        tb = torch.tensor(b, device="cuda:0", requires_grad=requires_grad)
        out.append((tw, tb))
    return out


def test_inference_matches_fp32_reference(ctx, batch):
    import torch

    noisy, albedo, gbuf, ref = batch
    dn = learned.Denoiser(ctx, seed=3)
    # Non-trivial last layer so every layer matters.
    w = dn.get_weights()
    rng = np.random.default_rng(0)
    last = sum(co * ci * 9 + co for co, ci in learned.layer_shapes()[:5])
    w[last:] = rng.normal(0, 0.05, len(w) - last).astype(np.float32)
    dn.set_weights(w)
    out = torch.empty_like(noisy)
    dn.infer(noisy, albedo, gbuf, out)
    torch.cuda.synchronize()
    ref_log = torch_reference(_weights_torch(dn), noisy, albedo, gbuf)
    got_log = torch.log1p(out[..., :3])
    err = (got_log - ref_log).abs().max().item()
    print(f"max |log out - log ref| = {err:.2e}")
    assert err < 2e-3
    assert torch.all(out[..., 3] == 1.0)
    # Without BF16 emulation the FP32 network differs by the BF16 rounding only.
    full = torch_reference(_weights_torch(dn), noisy, albedo, gbuf, emulate_bf16=False)
    assert (got_log - full).abs().max().item() < 0.1


def test_inference_at_ragged_sizes_matches_reference_and_training_forward(ctx):
    """At sizes that are not multiples of the 16 x 16 tile (partial tiles,
    image borders inside a tile's halo) inference matches the FP32
    reference, and its output gives the loss the training forward pass
    reports for the same input."""
    import torch

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    for (B, H, W) in ((1, 37, 53), (3, 17, 16), (2, 5, 70)):
        noisy = torch.tensor(rng.gamma(1.0, 0.5, (B, H, W, 4)).astype(np.float32), device=dev)
        albedo = torch.tensor(rng.random((B, H, W, 4)).astype(np.float32), device=dev)
        gbuf = torch.tensor(rng.normal(0, 1, (B, H, W, 4)).astype(np.float32), device=dev)
        gbuf[..., 3] = torch.tensor(rng.uniform(0.5, 30, (B, H, W)).astype(np.float32), device=dev)
        ref = torch.tensor(rng.gamma(1.0, 0.5, (B, H, W, 4)).astype(np.float32), device=dev)
        dn = learned.Denoiser(ctx, seed=11)
        w = dn.get_weights()
        last = sum(co * ci * 9 + co for co, ci in learned.layer_shapes()[:5])
        w[last:] = rng.normal(0, 0.05, len(w) - last).astype(np.float32)
        dn.set_weights(w)
        out = torch.empty_like(noisy)
        dn.infer(noisy, albedo, gbuf, out)
        torch.cuda.synchronize()
        ref_log = torch_reference(_weights_torch(dn), noisy, albedo, gbuf)
        got_log = torch.log1p(out[..., :3])
        err = (got_log - ref_log).abs().max().item()
        print(f"{B}x{H}x{W}: max |log out - log ref| = {err:.2e}")
        # Random inputs push some BF16 activations across a rounding boundary
        # between the FP32 accumulation orders of F.conv2d and the MFMAs; the
        # tolerance is wider than for rendered inputs.
        assert err < 1e-2
        # the training forward pass's loss on the same input
        _, loss_layers = dn.gradient(noisy, albedo, gbuf, ref)
        loss_fused = torch.mean((got_log - torch.log1p(torch.clamp(ref[..., :3], min=0.0))) ** 2).item()
        assert abs(loss_fused - loss_layers) <= 2e-4 * abs(loss_layers), (loss_fused, loss_layers)
        dn.close()


@pytest.mark.parametrize("kind", ["log_l2", "rel_mse"])
def test_gradient_matches_autograd(ctx, batch, kind):
    import torch

    noisy, albedo, gbuf, ref = batch
    dn = learned.Denoiser(ctx, seed=5)
    dn.set_loss(kind)
    w = dn.get_weights()
    rng = np.random.default_rng(1)
    last = sum(co * ci * 9 + co for co, ci in learned.layer_shapes()[:5])
    w[last:] = rng.normal(0, 0.05, len(w) - last).astype(np.float32)
    dn.set_weights(w)
    g, loss = dn.gradient(noisy, albedo, gbuf, ref)
    tw = _weights_torch(dn, requires_grad=True)
    out_log = torch_reference(tw, noisy, albedo, gbuf)
    r = torch.clamp(ref[..., :3], min=0.0)
    if kind == "log_l2":
        tl = torch.mean((out_log - torch.log1p(r)) ** 2)
    else:
        tl = torch.mean((torch.expm1(out_log) - r) ** 2 / (r * r + 1e-2))
    tl.backward()
    assert abs(loss - tl.item()) <= 1e-3 * abs(tl.item())
    k = 0
    for l, (W, b) in enumerate(tw):
        for t in (W, b):
            n = t.numel()
            a = g[k:k + n]
            r = t.grad.detach().cpu().numpy().ravel()
            k += n
            rel = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            cos = float(a @ r / max(np.linalg.norm(a) * np.linalg.norm(r), 1e-30))
            print(f"layer {l} {'W' if t is W else 'b'}: rel {rel:.2e} cos {cos:.5f}")
            assert rel < 5e-2 and cos > 0.998


def test_reference_render_is_the_mean_of_the_later_frames(ctx):
    # Training targets must be the plain mean of frames spp_in.. (the
    # renderer's running mean is a plain mean only from frame 0); the FP64
    # oracle window estimator gives that mean directly.
    import torch

    import oracle

    w = configs.c1(width=32, height=32)
    ctx.upload_scene(w.buffers)
    dev = torch.device("cuda:0")
    noisy = torch.empty((32, 32, 4), device=dev)
    gbuf, albedo, ref = torch.empty_like(noisy), torch.empty_like(noisy), torch.empty_like(noisy)
    s = torch.cuda.current_stream(dev).cuda_stream
    ctx.render_aux_dev(w.camera, 32, 32, 4, 0, 8, None, noisy.data_ptr(), gbuf.data_ptr(), albedo.data_ptr(), s)
    learned.reference_render(ctx, w.camera, 32, 32, 4, 16, 8, noisy, ref, gbuf, s)
    got = ref.cpu().numpy()[..., :3]
    expect = oracle.reference64_window(oracle.Scene(w.buffers), w.camera, 32, 32, 0, 0, 32, 32, 16, 4, 8)
    close = np.isclose(got, expect, rtol=1e-3, atol=1e-4).all(-1)
    print(f"pixels within 1e-3: {close.mean():.4f}")
    assert close.mean() >= 0.99


def test_committed_weights_reach_the_target_on_the_1080p_frame(ctx):
    # The trained weights shipped with the package, on the 1920 x 1080
    # icosphere frame at 8 spp: relMSE on the reference windows (65,536-spp
    # FP64 oracle) at most 1e-2, and below the bilateral filter's at the same
    # spp.  Rendering is deterministic, so this is a fixed-input check.
    import torch

    crops, origins, _ = learned.load_reference_crops()
    wl = learned.eval_workload()
    ctx.upload_scene(wl.buffers)
    dev = torch.device("cuda:0")
    W, H = wl.width, wl.height
    acc = torch.empty((H, W, 4), device=dev)
    gb, al, lo, bo = (torch.empty_like(acc) for _ in range(4))
    s = torch.cuda.current_stream(dev).cuda_stream
    ctx.render_aux_dev(wl.camera, W, H, 8, 0, wl.depth, None, acc.data_ptr(), gb.data_ptr(), al.data_ptr(), s)
    dn = learned.load_trained(ctx)
    dn.infer(acc[None], al[None], gb[None], lo[None], s)
    ctx.denoise_bilateral_dev(None, W, H, acc.data_ptr(), gb.data_ptr(), bo.data_ptr(), s)
    torch.cuda.synchronize()
    e_l = learned.crops_rel_mse(lo.cpu().numpy(), crops, origins)
    e_b = learned.crops_rel_mse(bo.cpu().numpy(), crops, origins)
    e_n = learned.crops_rel_mse(acc.cpu().numpy(), crops, origins)
    print(f"relMSE at 8 spp: noisy {e_n:.4e} bilateral {e_b:.4e} learned {e_l:.4e}")
    assert e_l <= 1e-2 and e_l < e_b < e_n


def test_nee_weights_reach_the_target_at_one_spp(ctx):
    # Weights trained on next-event-estimation renders, on a 1-spp NEE render
    # of the 1920 x 1080 icosphere frame: relMSE on the reference windows at
    # most 1e-2 (deterministic input).
    import torch

    crops, origins, _ = learned.load_reference_crops()
    wl = learned.eval_workload()
    ctx.upload_scene(wl.buffers)
    dev = torch.device("cuda:0")
    W, H = wl.width, wl.height
    acc = torch.empty((H, W, 4), device=dev)
    gb, al, lo = (torch.empty_like(acc) for _ in range(3))
    s = torch.cuda.current_stream(dev).cuda_stream
    ctx.set_nee(True)
    try:
        ctx.render_aux_dev(wl.camera, W, H, 1, 0, wl.depth, None, acc.data_ptr(), gb.data_ptr(), al.data_ptr(), s)
    finally:
        ctx.set_nee(False)
    dn = learned.load_trained(ctx, learned.WEIGHTS_NEE_PATH)
    dn.infer(acc[None], al[None], gb[None], lo[None], s)
    torch.cuda.synchronize()
    e = learned.crops_rel_mse(lo.cpu().numpy(), crops, origins)
    print(f"relMSE at 1 spp with NEE + learned: {e:.4e}")
    assert e <= 1e-2


def test_training_reduces_loss(ctx, batch):
    noisy, albedo, gbuf, ref = batch
    dn = learned.Denoiser(ctx, seed=9)
    losses = [dn.train_step(noisy, albedo, gbuf, ref, 2e-3) for _ in range(60)]
    print(f"train loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.5 * losses[0]
