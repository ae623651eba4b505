"""Timing DB and cost-model weights for the ragged path's strategy choice."""
