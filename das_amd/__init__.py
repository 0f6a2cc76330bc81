"""MI355X-native reverse-time-migration engine for the tsw_rtm (ll23002/DAS) RTM path.

Drop-in modules:
  das_amd.wave_solver   for tsw_rtm.wave_solver   (WaveSolver2D, ricker_wavelet)
  das_amd.rtm_engine    for tsw_rtm.rtm_engine    (TSWRTMEngine, RTMResult,
                                                   apply_laplacian_filter)
  das_amd.velocity_flood.FloodModels              (input container)
All compute runs in libtswrtm_hip.so (HIP, gfx950) through ctypes.
"""
__version__ = "0.1.0"
