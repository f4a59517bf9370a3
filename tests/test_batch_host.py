"""Host side of the batched-box workload: the std::mt19937_64 engine (C++11
[rand.predef] known answer) and the 48-box list it draws."""
from monofonic_amd.batch import MT19937_64, batch_boxes, uniform01


def test_mt19937_64_known_answer():
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_batch_box_list():
    boxes = batch_boxes()
    assert len(boxes) == 48
    assert [b[0] for b in boxes] == [128] * 16 + [192] * 16 + [256] * 16
    assert [b[1] for b in boxes] == list(range(12345, 12345 + 48))
    assert all(100.0 <= b[2] < 800.0 for b in boxes)
    assert sum(b[0] ** 3 for b in boxes) == 415236096
    g = MT19937_64(20261202)
    assert boxes[0][2] == 100.0 + 700.0 * (g() / 2.0 ** 64)
    u = [uniform01(g) for _ in range(1000)]
    assert 0.0 <= min(u) and max(u) < 1.0 and abs(sum(u) / 1000 - 0.5) < 0.05
