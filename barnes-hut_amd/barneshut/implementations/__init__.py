"""Implementations behind the ``[general] implementation`` key."""
