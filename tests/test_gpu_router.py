"""GPU tests of the learned router: its forward / backward / Adam kernels against the router
oracle (same bfloat16 operand rounding, float64 accumulation), training, and what the trained
router gains over ranking blocks without it."""
import numpy as np
import pytest

import knnhip
import oracle
import router_oracle as ro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# FP32 accumulation against the oracle's float64 (with identical bfloat16 operands): the gradient
# may differ by accumulation order and by the odd operand that lands on the other side of a
# bfloat16 rounding boundary; 2e-3 of the group's largest entry bounds both.
GRAD_TOL = 2e-3


def say(m):
    print(m, flush=True)


def groups(dim, H, nb):
    lay, _ = ro.layout(dim, H, nb)
    return {k: (o, int(np.prod(s))) for k, (o, s) in lay.items()}


@pytest.mark.parametrize("shape", [(24, 40, 70, 77, 5), (96, 256, 2048, 300, 16), (33, 64, 129, 1, 1)])
def test_loss_and_gradient_match_oracle(shape):
    dim, H, nb, n, k = shape
    say(f"router loss_grad dim={dim} H={H} nb={nb} n={n} k={k}")
    rng = np.random.default_rng(dim + H + nb)
    r = knnhip.Router(dim, H, nb, seed=11)
    params = r.get_params()
    # non-zero biases so that every term is exercised
    for name, (o, c) in groups(dim, H, nb).items():
        if name.startswith("b"):
            params[o:o + c] = rng.uniform(-0.2, 0.2, size=c).astype(np.float32)
    r.set_params(params)
    x = rng.normal(size=(dim, n)).astype(np.float32)
    tgt = rng.integers(0, nb, size=(k, n)).astype(np.int32)
    g, loss = r.loss_grad(torch.from_numpy(x).cuda(), torch.from_numpy(tgt).cuda())
    g = g.cpu().numpy().astype(np.float64)
    eloss, eg = ro.loss_grad(params, x, tgt, H, nb, rnd=True)
    assert abs(loss - eloss) <= 1e-5 * max(1.0, abs(eloss)), (loss, eloss)
    for name, (o, c) in groups(dim, H, nb).items():
        ref = eg[o:o + c]
        err = np.max(np.abs(g[o:o + c] - ref)) / (np.max(np.abs(ref)) + 1e-30)
        assert err < GRAD_TOL, f"{name}: relative error {err:.2e}"
    r.close()


def test_scores_match_oracle_forward():
    dim, H, nb, Q = 40, 64, 100, 130
    rng = np.random.default_rng(4)
    r = knnhip.Router(dim, H, nb, seed=5)
    params = r.get_params()
    x = rng.normal(size=(dim, Q)).astype(np.float32)
    s = r.scores(torch.from_numpy(x).cuda()).cpu().numpy()
    assert s.shape == (nb, Q)
    _, _, z3 = ro.forward(params, x, H, nb, rnd=True)
    err = np.max(np.abs(s - z3.T)) / np.max(np.abs(z3))
    assert err < 1e-4, err


def test_adam_step_matches_oracle():
    dim, H, nb, n, k = 20, 32, 50, 64, 4
    rng = np.random.default_rng(9)
    r = knnhip.Router(dim, H, nb, seed=3)
    p0 = r.get_params().astype(np.float64)
    x = torch.from_numpy(rng.normal(size=(dim, n)).astype(np.float32)).cuda()
    tgt = torch.from_numpy(rng.integers(0, nb, size=(k, n)).astype(np.int32)).cuda()
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0
    for step in (1, 2, 3):
        g, _ = r.loss_grad(x, tgt)  # the gradient the step will use (the kernels are deterministic)
        r.step(x, tgt, lr=1e-2)
        p, m, v = ro.adam(p, g.cpu().numpy().astype(np.float64), m, v, step, 1e-2)
        got = r.get_params().astype(np.float64)
        assert np.max(np.abs(got - p)) < 1e-5, f"step {step}: {np.max(np.abs(got - p)):.2e}"
        # keep the oracle on the GPU's FP32 parameters so that errors do not compound
        p = got
    r.close()


@pytest.fixture(scope="module")
def trained():
    # 64k refs (dim 32, 256 tight clusters) in 64 blocks; 8192 training queries from the mixture
    R, D, nb, k = 1 << 16, 32, 64, 8
    ref = knnhip.generate_mixture(R, D, 256, 0.05, 81, 82)
    train_q = knnhip.generate_mixture(8192, D, 256, 0.05, 81, 83)
    test_q = knnhip.generate_mixture(1024, D, 256, 0.05, 81, 84)
    ix = knnhip.BlockIndex(ref, nblocks=nb, iters=6)
    r = knnhip.Router(D, 128, nb, seed=7)
    loss0 = r.train(ix, train_q, k, epochs=1, batch=512, lr=1e-3)
    loss = r.train(ix, train_q, k, epochs=12, batch=512, lr=2e-3)
    say(f"router loss after 1 epoch {loss0:.4f}, after 13 epochs {loss:.4f}")
    yield ref, test_q, ix, r, k, nb, loss0, loss
    r.close()
    ix.close()


@pytest.mark.parametrize("shape", [(40, 64, 100, 300), (96, 256, 2048, 1000), (33, 128, 129, 64)])
def test_fused_top1_is_the_score_argmax(shape):
    # the fused first-probe kernel against the argmax of the score matrix: the two sum in a
    # different order (and a hidden unit near a bfloat16 rounding boundary may round the other
    # way), so the chosen block's score is within 4e-3 of the column's largest magnitude of the
    # best score, and the block is the argmax wherever the best two differ by more than 8e-3
    dim, H, nb, Q = shape
    say(f"fused top1 dim={dim} H={H} nb={nb} Q={Q}")
    rng = np.random.default_rng(dim + H)
    r = knnhip.Router(dim, H, nb, seed=3)
    params = r.get_params()
    for name, (o, c) in groups(dim, H, nb).items():
        if name.startswith("b"):
            params[o:o + c] = rng.uniform(-0.2, 0.2, size=c).astype(np.float32)
    r.set_params(params)
    x = torch.from_numpy(rng.normal(size=(dim, Q)).astype(np.float32)).cuda()
    top = r.top1(x).long()
    sc = r.scores(x)
    best = sc.max(dim=0)
    picked = sc.gather(0, top.unsqueeze(0)).squeeze(0)
    scale = sc.abs().max(dim=0).values
    assert bool(((best.values - picked) <= 4e-3 * scale).all())
    top2 = torch.topk(sc, 2, dim=0).values
    clear = (top2[0] - top2[1]) > 8e-3 * scale
    assert bool((top[clear] == best.indices[clear]).all())
    say(f"  argmax agreement {(top == best.indices).float().mean().item():.4f}")
    # parameters changed after the first call: the kernel's bfloat16 weights follow
    params[groups(dim, H, nb)["b3"][0]] += 100.0
    r.set_params(params)
    assert bool((r.top1(x) == 0).all())
    r.close()


def test_fused_top1_rejects_unsupported_hidden_width():
    r = knnhip.Router(24, 40, 70, seed=1)
    x = torch.zeros((24, 8), device="cuda")
    with pytest.raises(knnhip.KnnError, match="fused"):
        r.top1(x)
    r.close()


def test_training_reduces_loss(trained):
    ref, test_q, ix, r, k, nb, loss0, loss = trained
    assert loss < loss0
    assert loss < 0.5 * np.log(nb)


def neighbour_block_recall(scores, nbr_blocks, P):
    # share of each query's k neighbours that sit in its P top-scored blocks, averaged
    top = torch.topk(scores, P, dim=0).indices  # P x Q
    hit = (nbr_blocks.unsqueeze(1) == top.unsqueeze(0)).any(dim=1)  # k x Q
    return hit.float().mean().item()


def test_router_search_is_exact_and_gains_over_centroid_ranking(trained):
    ref, test_q, ix, r, k, nb, _, _ = trained
    pd, pi = knnhip.search_device(ref, test_q, k)
    start, count, ref_block = ix.layout()
    nbr_blocks = ref_block[pi.long()]  # k x Q
    scores = r.scores(test_q)
    # the method without the router, as an IVF index ranks: distance to the block centroids
    R = ref.shape[1]
    cent = torch.zeros((nb, ref.shape[0]), device="cuda")
    cent.index_add_(0, ref_block.long(), ref.t())
    cent /= torch.bincount(ref_block.long(), minlength=nb).clamp(min=1).unsqueeze(1).float()
    cscore = -torch.cdist(test_q.t().contiguous(), cent).t().contiguous()
    rec_r = neighbour_block_recall(scores, nbr_blocks, 1)
    rec_c = neighbour_block_recall(cscore, nbr_blocks, 1)
    say(f"neighbour share in the top block: router {rec_r:.4f}, centroid ranking {rec_c:.4f}")
    for P in (1, 2, 4):
        res = {}
        for name, sc, all_ in (("router", scores, True), ("centroid", cscore, True),
                               ("router first", scores, False), ("bounds", None, False)):
            d, i = ix.search(test_q, k, probes=P, scores=sc, score_all=all_)
            res[name] = knnhip.last_fallback_count()
            assert knnhip.last_plan()["scores_rank"] == ("none" if sc is None else "all" if all_ else "first")
            assert torch.equal(i, pi) and torch.equal(d, pd), f"{name}, probes {P}"
        say(f"probes {P}: certificate failures {res}")
    assert rec_r >= rec_c


def test_routed_search_is_exact(trained):
    # the router handle: its first probe from the fused kernel (or its scores ranking every
    # probe), the lower bounds for the rest; results equal the plain search
    ref, test_q, ix, r, k, nb, _, _ = trained
    pd, pi = knnhip.search_device(ref, test_q, k)
    for P in (1, 3):
        for prec in ("fp32", "bf16x3"):
            for all_ in (False, True):
                d, i = ix.search(test_q, k, probes=P, router=r, precision=prec, score_all=all_)
                plan = knnhip.last_plan()
                assert plan["router_kernel"] == ("scores" if all_ else "fused first probe")
                assert torch.equal(i, pi) and torch.equal(d, pd), f"probes {P} {prec} all={all_}"
    first = r.top1(test_q).long()
    start, count, ref_block = ix.layout()
    hit = (ref_block[pi[0].long()] == first).float().mean().item()
    say(f"nearest neighbour in the router's first probe: {hit:.4f}")
