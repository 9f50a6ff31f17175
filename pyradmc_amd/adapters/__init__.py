"""Input formats upstream of the transport path (pyradmc/adapters)."""
