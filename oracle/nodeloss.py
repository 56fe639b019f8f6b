"""CPU restatement of HierarchicalNeuralGeometry::AccumulateSampleLossPerNode (src/geometry/geometry.cpp:38-58) and
HyperTreeBaseImpl::SetErrorForActiveNodes (src/structure/HyperTree.cpp:225-265) -- TEST INFRASTRUCTURE ONLY: tests/
use it as the checker of neat_node_loss_accumulate and of the module layer; the product never imports it.

accumulate: per sample loss[ray] * weight (float64), summed per node, over the oracle sampler's emitted samples
(oracle.raysample with emit=True). Pinned: the weight sums equal the nodes' total segment lengths (the midpoint-cell
weights of a segment telescope to tmax - tmin).
set_error: the reference's tensor expressions in numpy."""
import numpy as np


def accumulate(samples, per_ray_loss, num_nodes):
    loss = np.asarray(per_ray_loss, np.float64)
    if loss.ndim == 2:
        loss = loss.mean(axis=0)
    w = samples["weight"].astype(np.float64)
    node = samples["node_id"].astype(np.int64)
    ls = np.zeros(num_nodes)
    ws = np.zeros(num_nodes)
    np.add.at(ls, node, loss[samples["ray_index"]] * w)
    np.add.at(ws, node, w)
    return ls, ws


def set_error(node_error, node_active, error, strategy):
    active = node_active.astype(np.float32)
    inactive = 1 - active
    invalid = (node_error < 0).astype(np.float32)
    error = active * error
    if strategy == "override":
        return node_error * inactive + error
    if strategy == "min":
        error = error + inactive * 1000000
        return np.minimum(node_error + (invalid * active) * 1000000, error)
    raise ValueError(strategy)
