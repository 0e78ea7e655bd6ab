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
