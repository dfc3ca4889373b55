"""orihime_amd — soft Smith-Waterman for AMD Instinct MI355X (gfx950)."""

__version__ = "0.1.0"
