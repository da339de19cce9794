"""Learners on the MI355X pipeline (brax/training), in torch."""
