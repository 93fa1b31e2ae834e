"""Parameters and helpers."""
