"""The reference's training loop body run as written, with only its imports changed.

experiments/mnist/train_spike_count.py:143-175 (forward, count loss, `.get()` of pred / loss /
output counts into the monitors, backward, `cp.mean(g, axis=0)`, Adam, apply_deltas) runs for
three steps on the script's own network and settings (784-800-10, thresholds 0.2 / 1.3, buffers
30 / 30, simulation time 0.2, weights ~ U(-1, 1) from `cp.random.uniform`, targets 15 / 3,
Adam at 0.003), with `cp` = bats_amd.cupy_compat. The batch is 16 instead of 50 (the oracle's
time) and the data are synthetic images in Dataset.get_train_batch's format
(experiments/mnist/Dataset.py:66-71: latency times in a 100 ms window, +inf for dark pixels).

Between the loop's own lines the test checks each stage against the FP64 oracle on the
weights the GPU holds at the start of the step: spike counts identical and times within 1e-5
relative; loss and errors within 1e-5 relative on the GPU's output counts; predictions equal;
batch-mean gradients within 1e-4 |ref| + 1e-6 max|ref|; Adam deltas equal to the oracle's
Adam fed the same gradients within 1e-6 relative; weights after the update equal to the
float32 sum. After the steps, the train monitors' records equal the statistics of the oracle's
losses and predictions (loss within 1e-6 relative, accuracy and silent-label shares exact).
"""
import numpy as np
import pytest
import torch

from oracle import lif as olif
from oracle import losses as oloss
from oracle import optim as oopt
from tests.helpers import assert_close_rel, spike_slots

pytestmark = pytest.mark.gpu
F32 = np.float32

# train_spike_count.py:20-52
N_INPUTS = 28 * 28
SIMULATION_TIME = 0.2
N_NEURONS_1 = 800
TAU_S_1 = 0.130
THRESHOLD_HAT_1 = 0.2
DELTA_THRESHOLD_1 = 1 * THRESHOLD_HAT_1
SPIKE_BUFFER_SIZE_1 = 30
N_OUTPUTS = 10
TAU_S_OUTPUT = 0.130
THRESHOLD_HAT_OUTPUT = 1.3
DELTA_THRESHOLD_OUTPUT = 1 * THRESHOLD_HAT_OUTPUT
SPIKE_BUFFER_SIZE_OUTPUT = 30
TRAIN_BATCH_SIZE = 16
LEARNING_RATE = 0.003
TARGET_FALSE = 3
TARGET_TRUE = 15


class SyntheticDataset:
    """Batches in Dataset.get_train_batch's format: times [B, 784, 1] (float64, +inf for
    pixels at 0), counts [B, 784], labels [B]; images with about MNIST's share of lit pixels."""

    TIME_WINDOW, MAX_VALUE = 100e-3, 255

    def __init__(self, rng, n):
        img = rng.integers(1, 256, size=(n, N_INPUTS))
        img[rng.uniform(size=img.shape) > 0.19] = 0
        self.images, self.labels = img, rng.integers(0, 10, size=n)

    def get_train_batch(self, batch_index, batch_size):
        s = slice(batch_index * batch_size, (batch_index + 1) * batch_size)
        t = self.TIME_WINDOW * (1 - self.images[s].reshape(-1, N_INPUTS, 1) / self.MAX_VALUE)
        t[t == self.TIME_WINDOW] = np.inf
        n = np.isfinite(t).astype("int").reshape(-1, N_INPUTS)
        return t, n, self.labels[s]


def test_reference_training_loop_body_three_steps():
    import bats_amd.cupy_compat as cp
    from bats_amd.Monitors import (AccuracyMonitor, LossMonitor, MonitorsManager,
                                   SilentLabelsMonitor, TimeMonitor)
    from bats_amd.Layers import InputLayer, LIFLayer
    from bats_amd.Losses import SpikeCountClassLoss
    from bats_amd.Network import Network
    from bats_amd.Optimizers import AdamOptimizer

    # train_spike_count.py:60-61, 67-68 and 75-103, verbatim but for the monitors' export paths
    def weight_initializer(n_post: int, n_pre: int) -> cp.ndarray:
        return cp.random.uniform(-1.0, 1.0, size=(n_post, n_pre), dtype=cp.float32)

    cp.random.seed(1234)
    network = Network()
    input_layer = InputLayer(n_neurons=N_INPUTS, name="Input layer")
    network.add_layer(input_layer, input=True)
    hidden_layer = LIFLayer(previous_layer=input_layer, n_neurons=N_NEURONS_1, tau_s=TAU_S_1,
                            theta=THRESHOLD_HAT_1,
                            delta_theta=DELTA_THRESHOLD_1,
                            weight_initializer=weight_initializer,
                            max_n_spike=SPIKE_BUFFER_SIZE_1,
                            name="Hidden layer 1")
    network.add_layer(hidden_layer)
    output_layer = LIFLayer(previous_layer=hidden_layer, n_neurons=N_OUTPUTS, tau_s=TAU_S_OUTPUT,
                            theta=THRESHOLD_HAT_OUTPUT,
                            delta_theta=DELTA_THRESHOLD_OUTPUT,
                            weight_initializer=weight_initializer,
                            max_n_spike=SPIKE_BUFFER_SIZE_OUTPUT,
                            name="Output layer")
    network.add_layer(output_layer)
    loss_fct = SpikeCountClassLoss(target_false=TARGET_FALSE, target_true=TARGET_TRUE)
    optimizer = AdamOptimizer(learning_rate=LEARNING_RATE)
    train_loss_monitor = LossMonitor()
    train_accuracy_monitor = AccuracyMonitor()
    train_silent_label_monitor = SilentLabelsMonitor()
    train_time_monitor = TimeMonitor()
    train_monitors_manager = MonitorsManager([train_loss_monitor,
                                              train_accuracy_monitor,
                                              train_silent_label_monitor,
                                              train_time_monitor],
                                             print_prefix="Train | ")

    w0 = hidden_layer.weights.get()
    assert w0.dtype == F32 and w0.shape == (800, 784) and -1.0 <= w0.min() and w0.max() < 1.0
    dataset = SyntheticDataset(np.random.default_rng(3), 3 * TRAIN_BATCH_SIZE)
    ref_opt = oopt.Adam(LEARNING_RATE)
    lays = [hidden_layer, output_layer]
    thetas = [THRESHOLD_HAT_1, THRESHOLD_HAT_OUTPUT]
    Ks = [SPIKE_BUFFER_SIZE_1, SPIKE_BUFFER_SIZE_OUTPUT]
    all_losses, all_hits, all_silent = [], [], []
    train_time_monitor.start()
    for batch_idx in range(3):
        print(f"[reference loop] step {batch_idx}", flush=True)
        Ws = [l.weights.get() for l in lays]
        # ---- train_spike_count.py:143-175
        spikes, n_spikes, labels = dataset.get_train_batch(batch_idx, TRAIN_BATCH_SIZE)

        network.reset()
        network.forward(spikes, n_spikes, max_simulation=SIMULATION_TIME, training=True)
        out_spikes, n_out_spikes = network.output_spike_trains

        pred = loss_fct.predict(out_spikes, n_out_spikes)
        loss, errors = loss_fct.compute_loss_and_errors(out_spikes, n_out_spikes, labels)

        pred_cpu = pred.get()
        loss_cpu = loss.get()
        n_out_spikes_cpu = n_out_spikes.get()
        errors_cpu = errors.get()  # (check) before the backward's in-place reset chain

        train_loss_monitor.add(loss_cpu)
        train_accuracy_monitor.add(pred_cpu, labels)
        train_silent_label_monitor.add(n_out_spikes_cpu, labels)

        gradient = network.backward(errors)
        avg_gradient = [None if g is None else cp.mean(g, axis=0) for g, layer in zip(gradient, network.layers)]
        del gradient

        # (check) forward, loss and gradients against the oracle on this step's weights
        states, cur = [], spikes.astype(F32)
        for W, K, th in zip(Ws, Ks, thetas):
            s = olif.lif_forward(cur, W, TAU_S_1, th, th, K, tmax=SIMULATION_TIME)
            states.append(s)
            cur = s.t
        for k, (lay, s) in enumerate(zip(lays, states)):
            t_g, n_g = (v.get() for v in lay.spike_trains)
            assert np.array_equal(n_g, s.n), f"layer {k}: spike counts differ"
            m = spike_slots(s.n, t_g.shape[2])
            rel = np.abs(t_g[m].astype(np.float64) - s.t[m]) / np.abs(s.t[m].astype(np.float64))
            assert rel.size == 0 or rel.max() <= 1e-5, f"layer {k}: spike times"
        assert states[0].n.sum() > 0 and states[1].n.sum() > 0
        tg = oloss.count_targets(n_out_spikes_cpu, labels, TARGET_TRUE, TARGET_FALSE)
        ref_loss, ref_err = oloss.count_loss_and_errors(n_out_spikes_cpu, tg, Ks[-1])
        assert_close_rel(loss_cpu, ref_loss, 1e-5, 0.0, "loss")
        assert_close_rel(errors_cpu, ref_err, 1e-5, 1e-6, "loss errors")
        assert np.array_equal(pred_cpu, oloss.count_predict(n_out_spikes_cpu))
        e = ref_err
        for k in (1, 0):
            _, g, pe = olif.lif_backward(states[k], e, pre_trainable=k > 0)
            assert_close_rel(avg_gradient[k + 1].get(), olif.batch_mean(g), 1e-4, 1e-6,
                             f"layer {k} gradient")
            e = pe
        assert avg_gradient[0] is None
        g_used = [None if g is None else g.get() for g in avg_gradient]

        deltas = optimizer.step(avg_gradient)
        del avg_gradient

        network.apply_deltas(deltas)
        # (check) Adam against the oracle's fed the same gradients, and the update itself
        for k, (d, want) in enumerate(zip(deltas, ref_opt.step(g_used))):
            if want is None:
                assert d is None
                continue
            assert_close_rel(d.get(), want, 1e-6, 0.0, f"Adam delta {k}")
            assert np.array_equal(lays[k - 1].weights.get(), (Ws[k - 1] + d.get()).astype(F32))
        del deltas
        # ---- end of the loop body
        all_losses.append(ref_loss)
        all_hits.append(oloss.count_predict(n_out_spikes_cpu) == labels)
        all_silent.append(n_out_spikes_cpu[np.arange(len(labels)), labels] == 0)

    records = train_monitors_manager.record(0.1)
    train_monitors_manager.print(0.1)
    losses = np.concatenate(all_losses).astype(np.float64)
    assert_close_rel(np.array([records[train_loss_monitor]]),
                     np.array([losses[np.isfinite(losses)].mean()]), 1e-6, 0.0, "loss monitor")
    assert records[train_accuracy_monitor] == pytest.approx(
        100.0 * np.concatenate(all_hits).mean(), abs=1e-9)
    assert records[train_silent_label_monitor] == pytest.approx(
        100.0 * np.concatenate(all_silent).mean(), abs=1e-9)
    assert records[train_time_monitor] > 0.0
