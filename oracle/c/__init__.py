"""C/OpenMP part of the oracle (CPU baseline timing and mid-size checks)."""
