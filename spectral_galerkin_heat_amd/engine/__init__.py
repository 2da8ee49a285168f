"""Batched multi-cuboid engine and the request-serving loop."""
