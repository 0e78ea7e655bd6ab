"""Host logic of WarmStart.train (no GPU: the training step is replaced by a
scripted one, so only the loop's bookkeeping runs)."""
import numpy as np


def test_training_rolls_back_loss_spikes(geneig):
    """An epoch whose mean loss exceeds 1.5x the best so far returns to the best
    epoch's weights, Adam moments and step count, and halves the learning rate
    from then on; training ends on the best epoch's weights."""
    import torch

    L = geneig.learn
    m = L.WarmStart(seed=0, device="cpu")
    f0 = m.flat.clone()
    script = [1.0, 0.8, 0.6, 2.0, 0.7, 0.5, 0.9]   # epoch means: spikes at epochs 3 (> 0.9) and 6 (> 0.75)
    P, batch, lr, lr_final = 8, 4, 2e-3, 1e-4      # two steps per epoch
    rates = []

    def step(X, Cset, idx, rate, stream=None):
        rates.append(rate)
        m.flat += 1.0
        m.m += 1.0
        m.v += 2.0
        m.step += 1
        return torch.full((idx.numel(),), script[(len(rates) - 1) // 2], dtype=torch.float32)

    m.train_step = step
    X = torch.zeros((P, 8), dtype=torch.float32)
    hist = m.train(X, None, epochs=len(script), batch=batch, lr=lr, lr_final=lr_final, seed=0)
    assert np.allclose(hist, script)
    assert m.rollbacks == 2 and m.train_best == 0.5
    # best epoch 5: ten steps taken (epoch 3's two were undone by the first rollback)
    assert torch.allclose(m.flat, f0 + 10.0, rtol=0, atol=1e-5) and torch.equal(m.m, torch.full_like(m.m, 10.0))
    assert torch.equal(m.v, torch.full_like(m.v, 20.0)) and m.step == 10
    steps = len(script) * 2
    base = [lr * (lr_final / lr) ** (k / (steps - 1)) for k in range(steps)]
    factor = [1.0] * 8 + [0.5] * 6                  # halved after epoch 3; epoch 6's rollback comes last
    assert np.allclose(rates, [f * b for f, b in zip(factor, base)], rtol=1e-12)


def test_training_nan_epochs_never_become_best(geneig):
    """NaN losses (rank-deficient network outputs) are counted apart and never
    enter an epoch's mean.  An epoch with more than 1% NaN losses is a spike:
    it returns to the best weights, halves the learning rate and is never kept
    as the best, however low the mean of its finite losses."""
    import torch

    L = geneig.learn
    m = L.WarmStart(seed=0, device="cpu")
    f0 = m.flat.clone()
    # per-epoch sample losses (two steps of 4 samples): epoch 2 has half its
    # samples NaN with tiny finite losses beside them
    nan = float("nan")
    script = [[1.0] * 8, [0.8] * 8, [nan, 0.01] * 4, [0.7] * 8]
    calls = []

    def step(X, Cset, idx, rate, stream=None):
        calls.append(rate)
        m.flat += 1.0
        m.m += 1.0
        m.v += 2.0
        m.step += 1
        ep, half = divmod(len(calls) - 1, 2)
        return torch.tensor(script[ep][4 * half:4 * half + 4], dtype=torch.float32)

    m.train_step = step
    X = torch.zeros((8, 8), dtype=torch.float32)
    hist = m.train(X, None, epochs=4, batch=4, lr=2e-3, lr_final=1e-4, seed=0)
    assert m.nan_counts == [0, 0, 4, 0]
    assert np.allclose(hist, [1.0, 0.8, 0.01, 0.7])       # means over the finite losses
    assert m.rollbacks == 1 and abs(m.train_best - 0.7) < 1e-6
    # best = epoch 3 (after returning to epoch 1's state at step 4): 4 + 2 steps
    assert torch.allclose(m.flat, f0 + 6.0, rtol=0, atol=1e-5) and m.step == 6
    steps = 8
    base = [2e-3 * (1e-4 / 2e-3) ** (k / (steps - 1)) for k in range(steps)]
    assert np.allclose(calls, [b * f for b, f in zip(base, [1.0] * 6 + [0.5] * 2)], rtol=1e-12)
