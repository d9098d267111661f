"""Host-side admission of the serving loop (no GPU): FIFO with backfill,
rejection of requests larger than the whole pool on arrival, and an end
condition when requests wait with nothing running (the loop must not spin)."""
from alpaca_amd import gpu, serving
from alpaca_amd.serving import Admission, Request


def _req(rid, arrival, dim, nb):
    return Request(rid, arrival, dim, nb, 0.1, gpu.FP64, seed=rid)


def test_fifo_with_backfill_and_rejection():
    reqs = [_req(0, 0.0, 2, 32), _req(1, 0.1, 3, 8), _req(2, 0.2, 2, 4), _req(3, 0.3, 3, 20)]
    # pool of 600 3D blocks = 18000 2D blocks; request 3 (8000 3D blocks) never fits
    q = Admission(reqs, {3: 600, 2: 18000})
    q.arrive(0.25)
    assert [r.rid for r in q.waiting] == [0, 1, 2] and not q.rejected
    free = {"blocks": 1100}  # room for 1024 + 64 2D blocks, not for 512 3D blocks (15360)

    def fits(r):
        need = serving.req_blocks(r) * (30 if r.dim == 3 else 1)
        return need <= free["blocks"]

    admitted = []

    def admit(r):
        free["blocks"] -= serving.req_blocks(r) * (30 if r.dim == 3 else 1)
        admitted.append(r.rid)

    assert q.admit(fits, admit) == 2
    assert admitted == [0, 2]  # request 2 backfills past the 3D request 1
    assert [r.rid for r in q.waiting] == [1]
    q.arrive(1.0)
    assert [r.rid for r in q.rejected] == [3]
    assert q.next_arrival() is None


def test_unfittable_requests_do_not_spin():
    """A request that exceeds the pool is rejected at arrival, so with nothing
    live the loop's queues drain; and a waiting request that cannot be placed
    with nothing running ends the loop (stuck) instead of busy-waiting."""
    reqs = serving.generate_requests(rate=50.0, window=0.2, seed=1, sizes_2d=(32,),
                                     sizes_3d=(8,))
    q = Admission(reqs, {2: 723, 3: 24})  # 32^2 = 1024 2D blocks, 8^3 = 512 3D blocks
    q.arrive(10.0)
    assert len(q.rejected) == len(reqs) and not q.waiting
    assert not q.stuck(0)
    q2 = Admission([_req(0, 0.0, 2, 4)], {2: 100, 3: 3})
    q2.arrive(1.0)
    assert q2.admit(lambda r: False, lambda r: None) == 0  # e.g. no case slot free
    assert q2.stuck(0) and not q2.stuck(1)


def test_shortest_first_with_a_live_cell_cap():
    """order='shortest' admits the smallest estimated work first and stops at
    the live-cell cap (but never leaves the device idle); a request past
    max_wait goes first whatever its size."""
    big = Request(0, 0.0, 3, 8, 0.25, gpu.FP64, seed=0)     # 2M cells, long
    mid = Request(1, 0.1, 2, 32, 0.25, gpu.FP64, seed=1)    # 262K cells
    small = Request(2, 0.2, 2, 4, 0.05, gpu.FP64, seed=2)   # 4K cells, short
    assert serving.req_work(small) < serving.req_work(mid) < serving.req_work(big)
    q = Admission([big, mid, small], {2: 10 ** 6, 3: 10 ** 5}, order="shortest",
                  live_cells_cap=300_000)
    q.arrive(1.0)
    got = []
    assert q.admit(lambda r: True, lambda r: got.append(r.rid), live_cells=0, now=1.0) == 2
    assert got == [2, 1]  # the cap (300K cells) holds the 2M-cell request back
    assert [r.rid for r in q.waiting] == [0]
    # with nothing live the first candidate is admitted even above the cap
    assert q.admit(lambda r: True, lambda r: got.append(r.rid), live_cells=0, now=1.0) == 1
    assert got == [2, 1, 0]
    # aging: past max_wait the big request overtakes smaller ones
    q2 = Admission([big, small], {2: 10 ** 6, 3: 10 ** 5}, order="shortest",
                   live_cells_cap=10 ** 9, max_wait=5.0)
    q2.arrive(1.0)
    order = []
    q2.admit(lambda r: True, lambda r: order.append(r.rid), now=10.0)
    assert order == [0, 2]
