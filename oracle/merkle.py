"""Binary Merkle tree with level injections (numpy).

Follows internal/merkle/tree.go:117-368: nodes are 1-indexed (leaf i at
nodes[n+i], root at nodes[1]); after each level's 2-to-1 compression, every
injection whose LevelWidth equals that level's width folds into it in schedule
order, node_j <- H(node_j, inj[j]); an injection as wide as the leaves folds
into the leaves before the first compression (tree.go:210-296).  Openings
follow tree.go:341-368.  Pinned by the brute-force restatement in the tests
(the same recursion written node by node, as merkle/injection_test.go:37-58
does with ``referenceRoot``) and by path verification (tree.go:392-436).
"""
import numpy as np

from .poseidon2 import node_compress


class Tree:
    def __init__(self, leaves, injections=()):
        """leaves: (8, n) digests; injections: list of (8, w) digest arrays in
        non-increasing width order."""
        n = leaves.shape[1]
        assert n & (n - 1) == 0
        self.n = n
        self.injections = [np.asarray(i, dtype=np.uint64) for i in injections]
        prev = n
        for inj in self.injections:
            w = inj.shape[1]
            assert w & (w - 1) == 0 and w <= prev
            prev = w
        nodes = np.zeros((8, 2 * n), dtype=np.uint64)
        nodes[:, n:] = leaves
        nxt = 0

        def fold(width):
            nonlocal nxt
            while nxt < len(self.injections) and self.injections[nxt].shape[1] == width:
                nodes[:, width:2 * width] = node_compress(nodes[:, width:2 * width],
                                                          self.injections[nxt])
                nxt += 1

        fold(n)
        start = n >> 1
        while start >= 1:
            nodes[:, start:2 * start] = node_compress(nodes[:, 2 * start:4 * start:2],
                                                      nodes[:, 2 * start + 1:4 * start:2])
            fold(start)
            start >>= 1
        assert nxt == len(self.injections)
        self.nodes = nodes

    def root(self):
        return [int(v) for v in self.nodes[:, 1]]

    def open(self, idx):
        """Returns (leaf_idx, siblings [depth x 8], injection_leaves [k x 8])."""
        depth = self.n.bit_length() - 1
        sibs = []
        pos = self.n + idx
        for _ in range(depth):
            sibs.append([int(v) for v in self.nodes[:, pos ^ 1]])
            pos >>= 1
        injl = []
        for inj in self.injections:
            d = (self.n // inj.shape[1]).bit_length() - 1
            injl.append([int(v) for v in inj[:, idx >> d]])
        return idx, sibs, injl
