"""Shared builders for tests: a complete problem (bases, ClusterOps,
ClusteredOperator, flat arrays) from integrals and a cluster partition."""
import numpy as np

from fermicg_amd.cluster_ops import compute_cluster_eigenbasis, compute_cluster_ops, add_cmf_operators
from fermicg_amd.operator import extract_ClusteredTerms
from fermicg_amd.marshal import flatten_ops, flatten_ham
from fermicg_amd.state import TPSCIstate, fock_configs_within


class Problem:
    def __init__(self, ints, clusters, ref_fock, max_roots, max_transfer=100, rdm=None, cmf=True):
        self.ints = ints
        self.clusters = clusters
        self.ref = tuple(tuple(f) for f in ref_fock)
        kw = dict(rdm1a=rdm[0], rdm1b=rdm[1]) if rdm is not None else {}
        self.bases = compute_cluster_eigenbasis(ints, clusters, max_roots=max_roots, **kw)
        self.ops = compute_cluster_ops(self.bases, ints)
        if cmf:
            from fermicg_amd.cluster_ops import product_state_rdm1
            Da, Db = rdm if rdm is not None else product_state_rdm1(self.bases, ints.n_orb, self.ref)
            add_cmf_operators(self.ops, self.bases, ints, Da, Db)
        self.H = extract_ClusteredTerms(ints, clusters)
        self.focks = fock_configs_within(self.bases, self.ref, max_transfer)
        self.fops = flatten_ops(self.ops, [len(c) for c in clusters])
        self.fham = flatten_ham(self.H)

    def full_space(self, R=1):
        return TPSCIstate.full_blocks(self.bases, self.focks, R=R)
