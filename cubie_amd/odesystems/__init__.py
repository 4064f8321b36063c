"""Subpackage of cubie_amd."""
