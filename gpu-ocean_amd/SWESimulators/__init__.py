"""MI355X (gfx950) implementation of gpu-ocean's CDKLM16 simulator step.

Put the directory that contains this package (``gpu-ocean_amd``) on
``sys.path`` and import as the reference's callers do::

    from SWESimulators import Common, CDKLM16
"""
