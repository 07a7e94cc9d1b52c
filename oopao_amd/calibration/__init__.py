"""Calibration: interaction matrix and its SVD pseudo-inverse (calibration/*.py)."""
