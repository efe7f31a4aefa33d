! Test-only harness around the reference's three streamer drivers
! (src/streamer_2d.f90, src/streamer_3d.f90, src/streamer_cyl.f90), turned
! into the modules ref_drv_2d, ref_drv_3d and ref_drv_cyl by
! oracle/drv2mod.awk at build time (oracle/Makefile). It gives the tests the
! drivers' own box procedures (fld_from_pot, fluxes_koren, update_solution,
! average_dens, set_photoi_rate, set_ref_flags, get_max_dt, set_init_cond,
! compute_fld) on the harness trees of ref_harness_2 / ref_harness_3, and the
! drivers' own main loop (drv_main), so that the C restatement and the host
! run loop can be compared with them on the same inputs. Nothing in the
! product links or loads this.
!
! variant: 1 = streamer_2d, 2 = streamer_3d, 3 = streamer_cyl.

module ref_drivers
  use iso_c_binding
  use m_lookup_table
  use m_streamer
  use m_a2_t, only: a2_t
  use m_a3_t, only: a3_t
  use m_a2_utils, only: a2_loop_box, a2_loop_box_arg, a2_loop_boxes_arg
  use m_a3_utils, only: a3_loop_box, a3_loop_box_arg, a3_loop_boxes_arg
  use m_a2_mg, only: mg2_t, mg2_init_mg, mg2_auto_op, mg2_auto_corr, mg2_auto_gsrb
  use m_a3_mg, only: mg3_init_mg, mg3_auto_op, mg3_auto_corr, mg3_auto_gsrb
  use ref_harness_2, only: h2 => tree
  use ref_harness_3, only: h3 => tree
  use ref_drv_2d, only: d2_main => drv_main, d2_tree => tree, d2_mg => mg, &
       d2_fld_from_pot => fld_from_pot, d2_fluxes_koren => fluxes_koren, &
       d2_update_solution => update_solution, d2_average_dens => average_dens, &
       d2_photoi_rate => set_photoi_rate, d2_ref_flags => set_ref_flags, &
       d2_max_dt => get_max_dt, d2_init_cond => set_init_cond, &
       d2_compute_fld => compute_fld, d2_sides_bc => sides_bc_pot, &
       d2_steps => drv_steps, d2_cells => drv_cells
  use ref_drv_3d, only: d3_main => drv_main, d3_tree => tree, d3_mg => mg, &
       d3_fld_from_pot => fld_from_pot, d3_fluxes_koren => fluxes_koren, &
       d3_update_solution => update_solution, d3_average_dens => average_dens, &
       d3_photoi_rate => set_photoi_rate, d3_ref_flags => set_ref_flags, &
       d3_max_dt => get_max_dt, d3_init_cond => set_init_cond, &
       d3_compute_fld => compute_fld, d3_sides_bc => sides_bc_pot, &
       d3_steps => drv_steps, d3_cells => drv_cells
  use ref_drv_cyl, only: dc_main => drv_main, dc_tree => tree, dc_mg => mg, &
       dc_fld_from_pot => fld_from_pot, dc_fluxes_koren => fluxes_koren, &
       dc_update_solution => update_solution, dc_average_dens => average_dens, &
       dc_photoi_rate => set_photoi_rate, dc_ref_flags => set_ref_flags, &
       dc_max_dt => get_max_dt, dc_init_cond => set_init_cond, &
       dc_compute_fld => compute_fld, dc_sides_bc => sides_bc_pot, &
       dc_steps => drv_steps, dc_cells => drv_cells
  implicit none
  private

  integer, parameter :: dp = kind(0.0d0)

contains

  ! The ST_* settings the drivers read from their config files
  ! (ST_load_cfg, src/m_streamer.f90:366-397). d = domain_len, applied_fld,
  ! dt_max, end_time, dt_out, epsilon_diel, gas_pressure, fld_mod_t0,
  ! fld_sin_amplitude, fld_sin_freq, fld_lin_deriv, time;
  ! iv = box_size, steps_amr, photoi_enabled.
  subroutine h_config(d, iv) bind(C, name="refdrv_config")
    real(c_double) :: d(12)
    integer(c_int) :: iv(3)
    ST_domain_len        = d(1)
    ST_applied_fld       = d(2)
    ST_dt_max            = d(3)
    ST_end_time          = d(4)
    ST_dt_out            = d(5)
    ST_epsilon_diel      = d(6)
    ST_gas_pressure      = d(7)
    ST_fld_mod_t0        = d(8)
    ST_fld_sin_amplitude = d(9)
    ST_fld_sin_freq      = d(10)
    ST_fld_lin_deriv     = d(11)
    ST_time              = d(12)
    ST_box_size          = iv(1)
    ST_steps_amr         = iv(2)
    ST_photoi_enabled    = iv(3) /= 0
    ST_applied_voltage   = -ST_domain_len * ST_applied_fld
    ST_sim_name          = "ref"
    ST_output_dir        = "."
  end subroutine h_config

  ! ST_init_cond (what ST_get_init_cond, src/m_streamer.f90:235-271, reads):
  ! r0(dim, n), r1(dim, n) segment ends, peak densities, widths, falloffs
  subroutine h_init_cond(dim, n, bg, r0, r1, dens, width, falloff) &
       bind(C, name="refdrv_init_cond")
    integer(c_int), value :: dim, n
    real(c_double), value :: bg
    real(c_double)        :: r0(dim, n), r1(dim, n), dens(n), width(n)
    integer(c_int)        :: falloff(n)
    if (allocated(ST_init_cond%seed_r0)) then
       deallocate(ST_init_cond%seed_r0, ST_init_cond%seed_r1, &
            ST_init_cond%seed_dens, ST_init_cond%seed_width, &
            ST_init_cond%seed_falloff)
    end if
    ST_init_cond%bg_dens = bg
    ST_init_cond%n_cond  = n
    allocate(ST_init_cond%seed_r0(dim, n), ST_init_cond%seed_r1(dim, n))
    allocate(ST_init_cond%seed_dens(n), ST_init_cond%seed_width(n))
    allocate(ST_init_cond%seed_falloff(n))
    ST_init_cond%seed_r0      = r0
    ST_init_cond%seed_r1      = r1
    ST_init_cond%seed_dens    = dens
    ST_init_cond%seed_width   = width
    ST_init_cond%seed_falloff = falloff
  end subroutine h_init_cond

  ! ST_td_tbl as ST_load_transport_data builds it (src/m_streamer.f90:273-333)
  ! from input rows x(n) and the columns cols(n, 4) (mobility, diffusion,
  ! alpha, eta), with the scale factors at their default of 1
  subroutine h_table(n_rows, max_fld, n, x, cols) bind(C, name="refdrv_table")
    integer(c_int), value :: n_rows, n
    real(c_double), value :: max_fld
    real(c_double)        :: x(n), cols(n, n_var_td)
    integer               :: k
    ST_td_tbl = LT_create(0.0_dp, max_fld, n_rows, n_var_td)
    do k = 1, n_var_td
       call LT_set_col(ST_td_tbl, k, x, cols(:, k))
    end do
  end subroutine h_table

  function h_time(dt) result(t) bind(C, name="refdrv_time")
    real(c_double), intent(out) :: dt
    real(c_double)              :: t
    t  = ST_time
    dt = ST_dt
  end function h_time

  ! The driver's main program from init_tree to the end of its time loop
  ! (src/streamer_2d.f90:34-143, src/streamer_3d.f90:34-143,
  ! src/streamer_cyl.f90:37-156, output statements left out); the final tree
  ! then becomes the harness tree of its dimension, where the accessors of
  ! ref_harness_2 / ref_harness_3 read it.
  subroutine h_run(variant) bind(C, name="refdrv_run")
    integer(c_int), value :: variant
    select case (variant)
    case (1)
       call d2_main()
       h2 = d2_tree
    case (2)
       call d3_main()
       h3 = d3_tree
    case (3)
       call dc_main()
       h2 = dc_tree
    end select
  end subroutine h_run

  ! The time steps and leaf cell-steps (leaf cells summed over the steps)
  ! that the driver's main loop has run since the last reset; reset = 1 sets
  ! them to zero after reading
  subroutine h_counters(variant, steps, cells, reset) bind(C, name="refdrv_counters")
    integer(c_int), value :: variant, reset
    integer(c_int)        :: steps
    real(c_double)        :: cells
    select case (variant)
    case (1)
       steps = d2_steps
       cells = d2_cells
       if (reset /= 0) then
          d2_steps = 0
          d2_cells = 0
       end if
    case (2)
       steps = d3_steps
       cells = d3_cells
       if (reset /= 0) then
          d3_steps = 0
          d3_cells = 0
       end if
    case (3)
       steps = dc_steps
       cells = dc_cells
       if (reset /= 0) then
          dc_steps = 0
          dc_cells = 0
       end if
    end select
  end subroutine h_counters

  ! The multigrid set-up of the driver's main program (src/streamer_2d.f90:
  ! 36-52, src/streamer_3d.f90:36-52, src/streamer_cyl.f90:39-55), for calls
  ! of compute_fld on the harness tree
  subroutine h_mg_init(variant) bind(C, name="refdrv_mg_init")
    integer(c_int), value :: variant
    select case (variant)
    case (1)
       call mg2_setup(d2_mg, 8)
       d2_mg%sides_bc => d2_sides_bc
       call mg2_init_mg(d2_mg)
    case (2)
       d3_mg%i_phi        = i_phi
       d3_mg%i_tmp        = i_fld
       d3_mg%i_rhs        = i_rhs
       d3_mg%i_eps        = i_eps
       d3_mg%n_cycle_base = 8
       d3_mg%sides_bc    => d3_sides_bc
       d3_mg%box_op      => mg3_auto_op
       d3_mg%box_corr    => mg3_auto_corr
       d3_mg%box_gsrb    => mg3_auto_gsrb
       call mg3_init_mg(d3_mg)
    case (3)
       call mg2_setup(dc_mg, 4)
       dc_mg%sides_bc => dc_sides_bc
       call mg2_init_mg(dc_mg)
    end select
  end subroutine h_mg_init

  subroutine mg2_setup(mg, n_base)
    type(mg2_t), intent(inout) :: mg
    integer, intent(in)        :: n_base
    mg%i_phi        = i_phi
    mg%i_tmp        = i_fld
    mg%i_rhs        = i_rhs
    mg%i_eps        = i_eps
    mg%n_cycle_base = n_base
    mg%box_op      => mg2_auto_op
    mg%box_corr    => mg2_auto_corr
    mg%box_gsrb    => mg2_auto_gsrb
  end subroutine mg2_setup

  subroutine h_compute_fld(variant, n_cycles, no_guess) bind(C, name="refdrv_compute_fld")
    integer(c_int), value :: variant, n_cycles, no_guess
    select case (variant)
    case (1)
       call d2_compute_fld(h2, n_cycles, no_guess /= 0)
    case (2)
       call d3_compute_fld(h3, n_cycles, no_guess /= 0)
    case (3)
       call dc_compute_fld(h2, n_cycles, no_guess /= 0)
    end select
  end subroutine h_compute_fld

  ! a$D_loop_box(tree, fld_from_pot), as compute_fld calls it
  subroutine h_fld_from_pot(variant) bind(C, name="refdrv_fld_from_pot")
    integer(c_int), value :: variant
    select case (variant)
    case (1)
       call a2_loop_box(h2, d2_fld_from_pot)
    case (2)
       call a3_loop_box(h3, d3_fld_from_pot)
    case (3)
       call a2_loop_box(h2, dc_fld_from_pot)
    end select
  end subroutine h_fld_from_pot

  ! a$D_loop_boxes_arg(tree, fluxes_koren, [dt], .true.), as the time loop
  ! calls it
  subroutine h_fluxes_koren(variant, dt) bind(C, name="refdrv_fluxes_koren")
    integer(c_int), value :: variant
    real(c_double), value :: dt
    select case (variant)
    case (1)
       call a2_loop_boxes_arg(h2, d2_fluxes_koren, [dt], .true.)
    case (2)
       call a3_loop_boxes_arg(h3, d3_fluxes_koren, [dt], .true.)
    case (3)
       call a2_loop_boxes_arg(h2, dc_fluxes_koren, [dt], .true.)
    end select
  end subroutine h_fluxes_koren

  ! a$D_loop_box_arg(tree, update_solution, [dt], .true.)
  subroutine h_update_solution(variant, dt) bind(C, name="refdrv_update_solution")
    integer(c_int), value :: variant
    real(c_double), value :: dt
    select case (variant)
    case (1)
       call a2_loop_box_arg(h2, d2_update_solution, [dt], .true.)
    case (2)
       call a3_loop_box_arg(h3, d3_update_solution, [dt], .true.)
    case (3)
       call a2_loop_box_arg(h2, dc_update_solution, [dt], .true.)
    end select
  end subroutine h_update_solution

  ! a$D_loop_box(tree, average_dens)
  subroutine h_average_dens(variant) bind(C, name="refdrv_average_dens")
    integer(c_int), value :: variant
    select case (variant)
    case (1)
       call a2_loop_box(h2, d2_average_dens)
    case (2)
       call a3_loop_box(h3, d3_average_dens)
    case (3)
       call a2_loop_box(h2, dc_average_dens)
    end select
  end subroutine h_average_dens

  ! a$D_loop_box_arg(tree, set_photoi_rate, [coeff], .true.), as
  ! set_photoionization calls it
  subroutine h_photoi_rate(variant, coeff) bind(C, name="refdrv_photoi_rate")
    integer(c_int), value :: variant
    real(c_double), value :: coeff
    select case (variant)
    case (1)
       call a2_loop_box_arg(h2, d2_photoi_rate, [coeff], .true.)
    case (2)
       call a3_loop_box_arg(h3, d3_photoi_rate, [coeff], .true.)
    case (3)
       call a2_loop_box_arg(h2, dc_photoi_rate, [coeff], .true.)
    end select
  end subroutine h_photoi_rate

  ! a$D_loop_box(tree, set_init_cond)
  subroutine h_init_cond_boxes(variant) bind(C, name="refdrv_set_init_cond")
    integer(c_int), value :: variant
    select case (variant)
    case (1)
       call a2_loop_box(h2, d2_init_cond)
    case (2)
       call a3_loop_box(h3, d3_init_cond)
    case (3)
       call a2_loop_box(h2, dc_init_cond)
    end select
  end subroutine h_init_cond_boxes

  ! The flags set_ref_flags gives every box of levels 1..max_lvl, in the
  ! order consistent_ref_flags calls it (afivo/src/m_aX_core.f90:636-652),
  ! with flags it leaves unset mapped to keep as there (:662-664).
  subroutine h_ref_flags(variant, n, flags) bind(C, name="refdrv_ref_flags")
    integer(c_int), value :: variant, n
    integer(c_int)        :: flags(n)
    integer               :: fl(n), lvl, i, id
    fl = -huge(1)
    select case (variant)
    case (1, 3)
       do lvl = 1, h2%max_lvl
          do i = 1, size(h2%lvls(lvl)%ids)
             id = h2%lvls(lvl)%ids(i)
             if (variant == 1) then
                call d2_ref_flags(h2%boxes, id, fl)
             else
                call dc_ref_flags(h2%boxes, id, fl)
             end if
          end do
       end do
    case (2)
       do lvl = 1, h3%max_lvl
          do i = 1, size(h3%lvls(lvl)%ids)
             id = h3%lvls(lvl)%ids(i)
             call d3_ref_flags(h3%boxes, id, fl)
          end do
       end do
    end select
    where (fl > 1 .or. fl < -1) fl = 0
    flags = fl
  end subroutine h_ref_flags

  function h_max_dt(variant) result(dt) bind(C, name="refdrv_max_dt")
    integer(c_int), value :: variant
    real(c_double)        :: dt
    select case (variant)
    case (1)
       dt = d2_max_dt(h2)
    case (2)
       dt = d3_max_dt(h3)
    case default
       dt = dc_max_dt(h2)
    end select
  end function h_max_dt

end module ref_drivers
