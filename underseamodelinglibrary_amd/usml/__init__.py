"""Python mirror of the reference's usml::{types, ocean, waveq3d} interfaces."""
