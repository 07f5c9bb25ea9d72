"""MI355X (gfx950) implementation of MicroMagnum's effective-field and
Runge-Kutta hot path.

  micromagnum_amd.magneto  - drop-in counterparts of the reference's magneto
                             functions (exchange, llge, anisotropy, Zhang-Li,
                             demag tensor + FFT convolution, RK algebra)
  micromagnum_amd.solver   - fused device-resident stepper
  micromagnum_amd.omf      - OMF/OVF 1.0 field files
The compute entry points run only on a HIP GPU with libmmhip.so built
(micromagnum_amd.build.build()).
"""
__version__ = "0.1.0"
