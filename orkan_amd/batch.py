"""Batched circuit evaluation (filled in with the batched launcher)."""
EXTRA = {}
