"""runko.tools counterparts: the comm_mode and particle enums
(src/runko/bindings/pytools.c++:23-35, src/runko/communication_common.h:30-38)."""

import enum


class comm_mode(enum.IntEnum):
    emf_J = 0
    emf_E = 1
    emf_B = 2
    pic_particle = 3
    pic_particle_extra = 4
    emf_J_exchange = 6


class particle(enum.IntEnum):
    electron = 0
    ion = 1
    photon = 2


def _virtual_tile_sync_handshake_mode(mode):
    """Handshake mode before an external sync (communication_common.h:42-49).
    One process owns every tile, so no handshake is ever needed."""
    return None


class threeD:
    pass
