"""Tuning tables (gfx950.json) and their compiled form."""
