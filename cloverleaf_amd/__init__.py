"""cloverleaf_amd: MI355X-native Embedding Propagation training step.

A drop-in for the EP path of Refefer/cloverleaf (``EmbeddingPropagator.learn_features``),
with the step implemented as hand-written HIP kernels for gfx950 behind the C ABI of
``include/cloverleaf_ep.h``.
"""
from .api import (  # noqa: F401
    Distance,
    EdgeType,
    EmbeddingPropagator,
    EPLoss,
    FeatureAggregator,
    FeatureNamespace,
    FeatureSet,
    Graph,
    GraphBuilder,
    LossWeighting,
    NeighborhoodAligner,
    NodeEmbedder,
    NodeEmbeddings,
    Vocab,
)

__all__ = [
    "Distance", "EdgeType", "EmbeddingPropagator", "EPLoss", "FeatureAggregator", "FeatureNamespace",
    "FeatureSet", "Graph", "GraphBuilder", "LossWeighting", "NeighborhoodAligner", "NodeEmbedder",
    "NodeEmbeddings", "Vocab",
]
