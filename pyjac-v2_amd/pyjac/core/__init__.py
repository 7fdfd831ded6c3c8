"""Mechanism front-end: parser, rate tables, Jacobian sparsity, samplers."""
