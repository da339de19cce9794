"""Generalized-coordinates pipeline (brax.generalized on MI355X)."""
